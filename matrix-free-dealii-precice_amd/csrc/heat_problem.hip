// Heat time-step pieces around the operator: quadrature-point coordinates,
// the right-hand side with inhomogeneous Dirichlet data
// (LaplaceProblem::assemble_rhs, heat_transfer.h:577-655) and the L2 error
// against an exact solution (LaplaceProblem::compute_error, :815-848).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "cell_kernels.hpp"
#include "common.hpp"
#include "operator.hpp"

namespace mfg {

// Physical coordinates of the quadrature points: the trilinear map of the
// cell's vertices at the Gauss points, output in input cell order.
template <int P>
__global__ void quadrature_points_kernel(const double *__restrict__ vertices, const int32_t *__restrict__ perm,
                                         const double *__restrict__ points, int n_cells, double *__restrict__ xyz) {
  constexpr int N = P + 1, N3 = N * N * N;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)n_cells * N3) return;
  const int c = (int)(g / N3), pt = (int)(g - (int64_t)c * N3);
  const double xi[3] = {points[pt % N], points[(pt / N) % N], points[pt / (N * N)]};
  const double *X = vertices + (int64_t)c * 24;
  double x[3] = {0, 0, 0};
  for (int v = 0; v < 8; ++v) {
    double phi = 1;
    for (int d = 0; d < 3; ++d) phi *= ((v >> d) & 1) ? xi[d] : 1 - xi[d];
    for (int e = 0; e < 3; ++e) x[e] += phi * X[v * 3 + e];
  }
  const int64_t o = ((int64_t)perm[c] * N3 + pt) * 3;
  for (int e = 0; e < 3; ++e) xyz[o + e] = x[e];
}

struct HeatRhsArgs {
  const uint32_t *cell_nodes;
  const int32_t *perm;  // stored cell -> input cell
  int cell_begin, cell_end;
  const double *u_old, *u_dir, *f;  // f per (input cell, point) or null
  const double *shape, *colloc;
  double dt;
  double *rhs;
};

// One colour of cells: with u_old and the Dirichlet lift u_D read as stored
// (the right-hand side uses the operator without constraints), submit the
// value JxW (u_old + dt f - u_D) and the gradient of -dt c grad u_D, i.e.
// int (u_old + dt f) phi - A u_D.  Scatter as the operator does (the first
// colour touching a node stores).
template <int P, int CPB>
__global__ void __launch_bounds__((block_threads<P, CPB>())) heat_rhs_cells(HeatRhsArgs a, QHeat<double> qop) {
  constexpr int N = P + 1, N3 = N * N * N;
  __shared__ double s_shape[N * N], s_colloc[N * N];
  __shared__ double s_tile[CPB * 6 * N3];
  const int t = threadIdx.x;
  if (t < N * N) {
    s_shape[t] = a.shape[t];
    s_colloc[t] = a.colloc[t];
  }
  const int lc_raw = t / N3;
  const bool pad = lc_raw >= CPB;
  const int lc = pad ? CPB - 1 : lc_raw;
  const int pt = pad ? 0 : t - lc_raw * N3;
  Tile<N, double> T;
  T.base = s_tile + lc * 6 * N3;
  T.pt = pt;
  T.i = pt % N;
  T.j = (pt / N) % N;
  T.k = pt / (N * N);
  T.writer = !pad;
  const int cell = a.cell_begin + (int)blockIdx.x * CPB + lc;
  const bool active = !pad && cell < a.cell_end;
  const int64_t qidx = active ? (int64_t)cell * N3 + pt : 0;
  const uint32_t entry = active ? a.cell_nodes[qidx] : 0u;
  const int64_t node = entry & kNodeMaskDev;
  const double u2[2] = {active ? a.u_dir[node] : 0.0, active ? a.u_old[node] : 0.0};
  const typename QHeat<double>::Data qd = qop.load(qidx, active ? cell : 0, pt);
  const double f = (active && a.f) ? a.f[(int64_t)a.perm[cell] * N3 + pt] : 0.0;
  __syncthreads();
  double val[2], grad[2][3];
  evaluate_vg<N, 2, double>(T, s_shape, s_colloc, u2, val, grad);
  double v1[1] = {val[0]}, g1[1][3] = {{grad[0][0], grad[0][1], grad[0][2]}};
  qop.apply(qd, v1, g1);  // A u_D at the point: JxW u_D and dt c JxW J^-T J^-1 grad u_D
  double vs[1] = {qd.jxw * (val[1] + a.dt * f) - v1[0]};
  double gs[1][3] = {{-g1[0][0], -g1[0][1], -g1[0][2]}};
  double out[1];
  integrate_vg<N, 1, double>(T, s_shape, s_colloc, vs, gs, out);
  if (active) a.rhs[node] = ((entry >> 31) ? 0.0 : a.rhs[node]) + out[0];
}

__global__ void zero_rows_kernel(double *v, const int32_t *rows, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[rows[i]] = 0.0;
}

struct HeatErrorArgs {
  const uint32_t *cell_nodes;
  const int32_t *perm;
  int n_cells;
  const double *u, *exact;  // exact per (input cell, point)
  const double *shape, *colloc;
  double *partials;         // one per workgroup
};

// sum over the group's points of (exact - u_h)^2 JxW, u_h read as stored
template <int P, int CPB>
__global__ void __launch_bounds__((block_threads<P, CPB>())) heat_error_cells(HeatErrorArgs a, QHeat<double> qop) {
  constexpr int N = P + 1, N3 = N * N * N, NT = block_threads<P, CPB>();
  __shared__ double s_shape[N * N], s_colloc[N * N];
  __shared__ double s_tile[CPB * 3 * N3];
  __shared__ double s_red[NT];
  const int t = threadIdx.x;
  if (t < N * N) {
    s_shape[t] = a.shape[t];
    s_colloc[t] = a.colloc[t];
  }
  const int lc_raw = t / N3;
  const bool pad = lc_raw >= CPB;
  const int lc = pad ? CPB - 1 : lc_raw;
  const int pt = pad ? 0 : t - lc_raw * N3;
  Tile<N, double> T;
  T.base = s_tile + lc * 3 * N3;
  T.pt = pt;
  T.i = pt % N;
  T.j = (pt / N) % N;
  T.k = pt / (N * N);
  T.writer = !pad;
  const int cell = (int)blockIdx.x * CPB + lc;
  const bool active = !pad && cell < a.n_cells;
  const int64_t qidx = active ? (int64_t)cell * N3 + pt : 0;
  const int64_t node = active ? (int64_t)(a.cell_nodes[qidx] & kNodeMaskDev) : 0;
  const double u1[1] = {active ? a.u[node] : 0.0};
  const typename QHeat<double>::Data qd = qop.load(qidx, active ? cell : 0, pt);
  const double ex = active ? a.exact[(int64_t)a.perm[cell] * N3 + pt] : 0.0;
  __syncthreads();
  double val[1], grad[1][3];
  evaluate_vg<N, 1, double>(T, s_shape, s_colloc, u1, val, grad);
  const double e = ex - val[0];
  s_red[t] = active ? e * e * qd.jxw : 0.0;
  __syncthreads();
  // fold the block (any size) into 64 values, then a tree; fixed order
  __shared__ double s_fold[64];
  if (t < 64) {
    double acc = 0;
    for (int k = t; k < NT; k += 64) acc += s_red[k];
    s_fold[t] = acc;
  }
  __syncthreads();
  for (int s = 32; s > 0; s >>= 1) {
    if (t < s) s_fold[t] += s_fold[t + s];
    __syncthreads();
  }
  if (t == 0) a.partials[blockIdx.x] = s_fold[0];
}

// cells per workgroup of the heat set-up kernels
template <int P>
constexpr int heat_cpb() {
  return P == 1 ? 8 : P == 2 ? 4 : P <= 4 ? 2 : 1;
}

template <int P>
static void quadrature_points_t(const OperatorData &op, double *xyz) {
  constexpr int N3 = (P + 1) * (P + 1) * (P + 1);
  const int64_t n = (int64_t)op.n_cells * N3;
  quadrature_points_kernel<P><<<(int)((n + 255) / 256), 256, 0, op.stream>>>(
      op.vertices.as<double>(), op.perm_dev.as<int32_t>(), op.tab_points.as<double>(), op.n_cells, xyz);
  MFG_HIP(hipGetLastError());
}

static QHeat<double> heat_qop(const OperatorData &op) {
  return QHeat<double>{op.qd.as<double>(), op.nqp, op.dt, op.affine_ref ? op.cell_geo.as<double>() : nullptr,
                       op.wq3.as<double>()};
}

template <int P>
static void heat_rhs_t(const OperatorData &op, double *rhs, const double *u_old, const double *u_dir,
                       const double *f) {
  constexpr int CPB = heat_cpb<P>();
  HeatRhsArgs a;
  a.cell_nodes = op.cell_nodes.as<uint32_t>();
  a.perm = op.perm_dev.as<int32_t>();
  a.u_old = u_old;
  a.u_dir = u_dir;
  a.f = f;
  a.shape = op.tab_shape.as<double>();
  a.colloc = op.tab_colloc.as<double>();
  a.dt = op.dt;
  a.rhs = rhs;
  const QHeat<double> q = heat_qop(op);
  for (size_t col = 0; col + 1 < op.colour_start.size(); ++col) {
    a.cell_begin = op.colour_start[col];
    a.cell_end = op.colour_start[col + 1];
    const int groups = (a.cell_end - a.cell_begin + CPB - 1) / CPB;
    if (groups == 0) continue;
    heat_rhs_cells<P, CPB><<<groups, block_threads<P, CPB>(), 0, op.stream>>>(a, q);
    MFG_HIP(hipGetLastError());
  }
  if (op.n_constrained > 0) {
    zero_rows_kernel<<<(int)((op.n_constrained + 255) / 256), 256, 0, op.stream>>>(rhs, op.constrained.as<int32_t>(),
                                                                                  op.n_constrained);
    MFG_HIP(hipGetLastError());
  }
}

template <int P>
static double heat_error_t(const OperatorData &op, const double *u, const double *exact) {
  constexpr int CPB = heat_cpb<P>();
  const int groups = (op.n_cells + CPB - 1) / CPB;
  DeviceBuffer partials((size_t)groups * sizeof(double));
  HeatErrorArgs a;
  a.cell_nodes = op.cell_nodes.as<uint32_t>();
  a.perm = op.perm_dev.as<int32_t>();
  a.n_cells = op.n_cells;
  a.u = u;
  a.exact = exact;
  a.shape = op.tab_shape.as<double>();
  a.colloc = op.tab_colloc.as<double>();
  a.partials = partials.as<double>();
  heat_error_cells<P, CPB><<<groups, block_threads<P, CPB>(), 0, op.stream>>>(a, heat_qop(op));
  MFG_HIP(hipGetLastError());
  std::vector<double> h((size_t)groups);
  MFG_HIP(hipMemcpyAsync(h.data(), partials.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost, op.stream));
  MFG_HIP(hipStreamSynchronize(op.stream));
  double s = 0;  // fixed order: deterministic
  for (double v : h) s += v;
  return std::sqrt(s);
}

#define MFG_DEGREE_DISPATCH(p, FN, ...)                                   \
  switch (p) {                                                            \
    case 1: return FN<1>(__VA_ARGS__);                                    \
    case 2: return FN<2>(__VA_ARGS__);                                    \
    case 3: return FN<3>(__VA_ARGS__);                                    \
    case 4: return FN<4>(__VA_ARGS__);                                    \
    case 5: return FN<5>(__VA_ARGS__);                                    \
    case 6: return FN<6>(__VA_ARGS__);                                    \
    default: throw Error(MFG_ERR_ARGUMENT, "unsupported degree");         \
  }

static void quadrature_points(const OperatorData &op, double *xyz) { MFG_DEGREE_DISPATCH(op.p, quadrature_points_t, op, xyz); }
static void heat_rhs(const OperatorData &op, double *rhs, const double *u_old, const double *u_dir, const double *f) {
  MFG_DEGREE_DISPATCH(op.p, heat_rhs_t, op, rhs, u_old, u_dir, f);
}
static double heat_error(const OperatorData &op, const double *u, const double *exact) {
  MFG_DEGREE_DISPATCH(op.p, heat_error_t, op, u, exact);
}

}  // namespace mfg

using namespace mfg;

extern "C" {

int mfg_operator_quadrature_points(mfg_operator op, double *xyz) {
  return guarded([&] {
    MFG_REQUIRE(op && xyz, MFG_ERR_ARGUMENT, "null argument");
    quadrature_points(*op, xyz);
  });
}

int mfg_heat_assemble_rhs(mfg_operator op, double *rhs, const double *u_old, const double *u_dirichlet,
                          const double *f_qp) {
  return guarded([&] {
    MFG_REQUIRE(op && rhs && u_old && u_dirichlet, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(op->kind == 1, MFG_ERR_ARGUMENT, "not a heat operator");
    MFG_REQUIRE(op->precision == MFG_FP64, MFG_ERR_ARGUMENT, "the right-hand side is assembled in FP64");
    MFG_REQUIRE(rhs != u_old && rhs != u_dirichlet, MFG_ERR_ARGUMENT, "rhs aliases an input");
    heat_rhs(*op, rhs, u_old, u_dirichlet, f_qp);
  });
}

int mfg_heat_l2_error(mfg_operator op, const double *u, const double *exact_qp, double *error) {
  return guarded([&] {
    MFG_REQUIRE(op && u && exact_qp && error, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(op->kind == 1, MFG_ERR_ARGUMENT, "not a heat operator");
    MFG_REQUIRE(op->precision == MFG_FP64, MFG_ERR_ARGUMENT, "the error is computed in FP64");
    *error = heat_error(*op, u, exact_qp);
  });
}

}  // extern "C"
