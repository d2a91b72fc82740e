// Newton residual (Solid::assemble_residual, fast FEEvaluation path,
// mf_elasticity.h:1764-1846): R = -int tau F^{-T} : Grad w dX
// + int rho (b - a) . w dX, on the reference configuration (MappingQ1).
#pragma once
#include "setup_kernels.hpp"
#include "vector_kernels.hpp"

namespace mfg {

template <typename Number>
struct ResidualArgs {
  const uint32_t *cell_nodes;
  int cell_begin, cell_end;
  const Number *u, *acc;
  Number *R;
  const double *vertices;
  const Number *shape, *colloc;
  const double *points, *weights;
  Number mu, lambda, kappa, c_1, rho;
  Number body[3];
  int formulation;
  int *error;
};

// values only (3 sweeps), slots 0..2NC-1
template <int N, int NC, typename Number>
__device__ inline void evaluate_v(const Tile<N, Number> &T, const Number *S, const Number (&u)[NC], Number (&val)[NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) T.put(c, u[c]);
  __syncthreads();
  Number t[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) t[c] = T.template line<0>(S, T.i, c);
#pragma unroll
  for (int c = 0; c < NC; ++c) T.put(NC + c, t[c]);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; ++c) t[c] = T.template line<1>(S, T.j, NC + c);
#pragma unroll
  for (int c = 0; c < NC; ++c) T.put(c, t[c]);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; ++c) val[c] = T.template line<2>(S, T.k, c);
}

template <int P, int CPB, typename Number>
__global__ void __launch_bounds__((block_threads<P, CPB>())) residual_cells(ResidualArgs<Number> a) {
  constexpr int N = P + 1, N3 = N * N * N;
  __shared__ Number s_shape[N * N], s_colloc[N * N];
  __shared__ Number s_tile[CPB * 9 * N3];
  const int t = threadIdx.x;
  if (t < N * N) {
    s_shape[t] = a.shape[t];
    s_colloc[t] = a.colloc[t];
  }
  const int lc_raw = t / N3;
  const bool pad = lc_raw >= CPB;
  const int lc = pad ? CPB - 1 : lc_raw;
  const int pt = pad ? 0 : t - lc_raw * N3;
  Tile<N, Number> T;
  T.base = s_tile + lc * 9 * N3;
  T.pt = pt;
  T.i = pt % N;
  T.j = (pt / N) % N;
  T.k = pt / (N * N);
  T.writer = !pad;
  const int cell = a.cell_begin + blockIdx.x * CPB + lc;
  const bool active = !pad && cell < a.cell_end;
  const int64_t qidx = active ? (int64_t)cell * N3 + pt : 0;
  const uint32_t entry = active ? a.cell_nodes[qidx] : 0u;
  const int64_t node = entry & kNodeMaskDev;
  Number u[3], ac[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {  // read_dof_values_plain
    u[c] = active ? a.u[node * 3 + c] : Number(0);
    ac[c] = active ? a.acc[node * 3 + c] : Number(0);
  }
  __syncthreads();
  Number av[3];
  evaluate_v<N, 3, Number>(T, s_shape, ac, av);
  __syncthreads();
  Number uval[3], gr[3][3];
  evaluate_vg<N, 3, Number>(T, s_shape, s_colloc, u, uval, gr);

  Number val[3], grad[3][3];
  {
    double Jr[3][3], Jri[3][3];
    trilinear_jacobian(a.vertices + (int64_t)(active ? cell : 0) * 24, a.points[T.i], a.points[T.j], a.points[T.k], Jr);
    const double det = Jr[0][0] * (Jr[1][1] * Jr[2][2] - Jr[1][2] * Jr[2][1]) -
                       Jr[0][1] * (Jr[1][0] * Jr[2][2] - Jr[1][2] * Jr[2][0]) +
                       Jr[0][2] * (Jr[1][0] * Jr[2][1] - Jr[1][1] * Jr[2][0]);
    const double id = 1.0 / det;
    Jri[0][0] = (Jr[1][1] * Jr[2][2] - Jr[1][2] * Jr[2][1]) * id;
    Jri[0][1] = (Jr[0][2] * Jr[2][1] - Jr[0][1] * Jr[2][2]) * id;
    Jri[0][2] = (Jr[0][1] * Jr[1][2] - Jr[0][2] * Jr[1][1]) * id;
    Jri[1][0] = (Jr[1][2] * Jr[2][0] - Jr[1][0] * Jr[2][2]) * id;
    Jri[1][1] = (Jr[0][0] * Jr[2][2] - Jr[0][2] * Jr[2][0]) * id;
    Jri[1][2] = (Jr[0][2] * Jr[1][0] - Jr[0][0] * Jr[1][2]) * id;
    Jri[2][0] = (Jr[1][0] * Jr[2][1] - Jr[1][1] * Jr[2][0]) * id;
    Jri[2][1] = (Jr[0][1] * Jr[2][0] - Jr[0][0] * Jr[2][1]) * id;
    Jri[2][2] = (Jr[0][0] * Jr[1][1] - Jr[0][1] * Jr[1][0]) * id;
    const Number jxw = (Number)(fabs(det) * a.weights[T.i] * a.weights[T.j] * a.weights[T.k]);
    Number F[3][3], Fi[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int e = 0; e < 3; ++e)
        F[d][e] = gr[d][0] * (Number)Jri[0][e] + gr[d][1] * (Number)Jri[1][e] + gr[d][2] * (Number)Jri[2][e] +
                  (d == e ? Number(1) : Number(0));
    const Number J = det3(F);
    if (active && !(J > Number(0))) atomicOr(a.error, 1);
    inv3(F, Fi);
    Number b[3][3], tau[3][3];
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
      for (int y = 0; y < 3; ++y) b[x][y] = F[x][0] * F[y][0] + F[x][1] * F[y][1] + F[x][2] * F[y][2];
    if (a.formulation == 1) {  // material.h:148-158
      const Number tmp = a.mu - Number(2) * a.lambda * log(J);
#pragma unroll
      for (int x = 0; x < 3; ++x)
#pragma unroll
        for (int y = 0; y < 3; ++y) tau[x][y] = a.mu * b[x][y] - (x == y ? tmp : Number(0));
    } else {  // material.h:125-147
      const Number Jm23 = pow(J, Number(-2.0 / 3.0));
      const Number dpsi = (a.kappa / Number(2)) * (J - Number(1) / J);
      const Number tr = Number(2) * a.c_1 * Jm23 * (b[0][0] + b[1][1] + b[2][2]);
#pragma unroll
      for (int x = 0; x < 3; ++x)
#pragma unroll
        for (int y = 0; y < 3; ++y)
          tau[x][y] = Number(2) * a.c_1 * Jm23 * b[x][y] + (x == y ? dpsi * J - tr / Number(3) : Number(0));
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      Number Pd[3];
#pragma unroll
      for (int e = 0; e < 3; ++e) Pd[e] = -(tau[d][0] * Fi[e][0] + tau[d][1] * Fi[e][1] + tau[d][2] * Fi[e][2]);
#pragma unroll
      for (int f = 0; f < 3; ++f)
        grad[d][f] = jxw * (Pd[0] * (Number)Jri[f][0] + Pd[1] * (Number)Jri[f][1] + Pd[2] * (Number)Jri[f][2]);
      val[d] = jxw * a.rho * (a.body[d] - av[d]);
    }
  }
  Number out[3];
  integrate_vg<N, 3, Number>(T, s_shape, s_colloc, val, grad, out);
  if (active) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (!((entry >> (kConstraintShift + c)) & 1u)) a.R[node * 3 + c] += out[c];
  }
}

// Newmark: a = a1 du - a2 v_old - a3 a_old (update_acceleration, :1614-1620)
template <typename T>
__global__ void newmark_acc_kernel(T *acc, const T *du, const T *v_old, const T *a_old, double a1, double a2,
                                   double a3, int64_t n) {
  MFG_GRID_STRIDE(i, n) acc[i] = (T)(a1 * (double)du[i] - a2 * (double)v_old[i] - a3 * (double)a_old[i]);
}
// v = a4 du + a5 v_old + a6 a_old (update_velocity, :1600-1609)
template <typename T>
__global__ void newmark_vel_kernel(T *vel, const T *du, const T *v_old, const T *a_old, double a4, double a5,
                                   double a6, int64_t n) {
  MFG_GRID_STRIDE(i, n) vel[i] = (T)(a4 * (double)du[i] + a5 * (double)v_old[i] + a6 * (double)a_old[i]);
}
// x += y ; z += y
template <typename T>
__global__ void add2_kernel(T *x, T *z, const T *y, int64_t n) {
  MFG_GRID_STRIDE(i, n) {
    x[i] += y[i];
    z[i] += y[i];
  }
}

}  // namespace mfg
