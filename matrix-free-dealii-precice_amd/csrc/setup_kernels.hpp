// Per-linearisation set-up kernels: the quadrature cache (NeoHookOperator::
// cache, mf_nh_operator.h:392-523, with the deformed-frame geometry that
// MappingQEulerian provides to the tensor2 caching, mf_elasticity.h:951-952)
// and the operator diagonal (compute_diagonal, :1023-1060).
#pragma once
#include "cell_kernels.hpp"

namespace mfg {

template <typename Number>
struct CacheArgs {
  const uint32_t *cell_nodes;
  int cell_begin, cell_end;  // stored cells of one material
  const void *u;     // displacement, precision u_fp32 ? float : double
  int u_fp32;
  const double *vertices;  // 24 per stored cell
  const Number *shape, *colloc;
  const double *points, *weights;  // Gauss rule on [0,1], N each
  Number *qd;
  int64_t nqp;
  Number mu, lambda;
  int formulation;
  int caching;  // MFG_CACHING_*
  int jxw_ref_field;  // tensor2 formulation 1: store JxW_ref as field 16 (not supplied per cell)
  int t2_gradient;    // tensor2 formulation 1: store Grad_xi u and s (10 fields) instead
  int *error;   // set to 1 if det F <= 0 anywhere
};

template <typename Number>
__device__ inline Number det3(const Number (&A)[3][3]) {
  return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
         A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}
template <typename Number>
__device__ inline void inv3(const Number (&A)[3][3], Number (&R)[3][3]) {
  const Number id = Number(1) / det3(A);
  R[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * id;
  R[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  R[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  R[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * id;
  R[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  R[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  R[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * id;
  R[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  R[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
}

// Trilinear Jacobian dX_e/dxi_f at reference point xi from 8 vertices.
__device__ inline void trilinear_jacobian(const double *v, double x0, double x1, double x2, double (&J)[3][3]) {
#pragma unroll
  for (int e = 0; e < 3; ++e)
#pragma unroll
    for (int f = 0; f < 3; ++f) J[e][f] = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int b0 = c & 1, b1 = (c >> 1) & 1, b2 = (c >> 2) & 1;
    const double n0 = b0 ? x0 : 1 - x0, n1 = b1 ? x1 : 1 - x1, n2 = b2 ? x2 : 1 - x2;
    const double d0 = b0 ? 1 : -1, d1 = b1 ? 1 : -1, d2 = b2 ? 1 : -1;
    const double g0 = d0 * n1 * n2, g1 = n0 * d1 * n2, g2 = n0 * n1 * d2;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      const double X = v[c * 3 + e];
      J[e][0] += X * g0;
      J[e][1] += X * g1;
      J[e][2] += X * g2;
    }
  }
}

template <int P, int CPB, typename Number>
__global__ void __launch_bounds__((block_threads<P, CPB>())) cache_solid(CacheArgs<Number> a) {
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
  const int64_t node = active ? (int64_t)(a.cell_nodes[qidx] & kNodeMaskDev) : 0;
  Number u[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    // read_dof_values_plain: constrained entries are read as stored
    if (!active) u[c] = 0;
    else if (a.u_fp32) u[c] = (Number)((const float *)a.u)[node * 3 + c];
    else u[c] = (Number)((const double *)a.u)[node * 3 + c];
  }
  __syncthreads();
  Number gr[3][3];
  evaluate_g<N, 3, Number>(T, s_shape, s_colloc, u, gr);
  if (!active) return;

  double Jr[3][3], Jri[3][3];
  trilinear_jacobian(a.vertices + (int64_t)cell * 24, a.points[T.i], a.points[T.j], a.points[T.k], Jr);
  {
    const double id = 1.0 / (Jr[0][0] * (Jr[1][1] * Jr[2][2] - Jr[1][2] * Jr[2][1]) -
                             Jr[0][1] * (Jr[1][0] * Jr[2][2] - Jr[1][2] * Jr[2][0]) +
                             Jr[0][2] * (Jr[1][0] * Jr[2][1] - Jr[1][1] * Jr[2][0]));
    Jri[0][0] = (Jr[1][1] * Jr[2][2] - Jr[1][2] * Jr[2][1]) * id;
    Jri[0][1] = (Jr[0][2] * Jr[2][1] - Jr[0][1] * Jr[2][2]) * id;
    Jri[0][2] = (Jr[0][1] * Jr[1][2] - Jr[0][2] * Jr[1][1]) * id;
    Jri[1][0] = (Jr[1][2] * Jr[2][0] - Jr[1][0] * Jr[2][2]) * id;
    Jri[1][1] = (Jr[0][0] * Jr[2][2] - Jr[0][2] * Jr[2][0]) * id;
    Jri[1][2] = (Jr[0][2] * Jr[1][0] - Jr[0][0] * Jr[1][2]) * id;
    Jri[2][0] = (Jr[1][0] * Jr[2][1] - Jr[1][1] * Jr[2][0]) * id;
    Jri[2][1] = (Jr[0][1] * Jr[2][0] - Jr[0][0] * Jr[2][1]) * id;
    Jri[2][2] = (Jr[0][0] * Jr[1][1] - Jr[0][1] * Jr[1][0]) * id;
  }
  // F = I + Grad_X u ; current-frame Jacobian dx/dxi = dX/dxi + du/dxi
  Number F[3][3], Jc[3][3], Jci[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      F[d][e] = gr[d][0] * (Number)Jri[0][e] + gr[d][1] * (Number)Jri[1][e] + gr[d][2] * (Number)Jri[2][e] +
                (d == e ? Number(1) : Number(0));
      Jc[d][e] = (Number)Jr[d][e] + gr[d][e];
    }
  const Number J = det3(F);
  if (!(J > Number(0))) {
    atomicOr(a.error, 1);
    return;
  }
  const int64_t n = a.nqp;
  const double wq = a.weights[T.i] * a.weights[T.j] * a.weights[T.k];
  const Number scalar = a.mu - Number(2) * a.lambda * log(J);
  if (a.t2_gradient) {  // the 10-word tensor2 storage: Grad_xi u [d][f], s, in pairs of fields
#pragma unroll
    for (int f = 0; f < 9; ++f) a.qd[((f >> 1) * n + qidx) * 2 + (f & 1)] = gr[f / 3][f % 3];
    a.qd[(4 * n + qidx) * 2 + 1] = scalar;
    return;
  }
  if (a.caching == MFG_CACHING_SCALAR_REFERENTIAL || a.caching == MFG_CACHING_TENSOR4_NS) {
    // reference-frame geometry (MappingQ1 of the reference configuration)
    const double detr = Jr[0][0] * (Jr[1][1] * Jr[2][2] - Jr[1][2] * Jr[2][1]) -
                        Jr[0][1] * (Jr[1][0] * Jr[2][2] - Jr[1][2] * Jr[2][0]) +
                        Jr[0][2] * (Jr[1][0] * Jr[2][1] - Jr[1][1] * Jr[2][0]);
#pragma unroll
    for (int f = 0; f < 9; ++f) a.qd[f * n + qidx] = (Number)Jri[f / 3][f % 3];
    a.qd[9 * n + qidx] = (Number)(wq * fabs(detr));
    if (a.caching == MFG_CACHING_SCALAR_REFERENTIAL) {  // mf_nh_operator.h:442-455
      a.qd[10 * n + qidx] = scalar;
      return;
    }
    // tensor4_ns (mf_nh_operator.h:495-516): 2 lambda F^-T (x) F^-T + s_c F^-T (x)_il F^-1 + mu I (x)_ik I
    Number Fi[3][3];
    inv3(F, Fi);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int l = 0; l < 3; ++l)
            a.qd[(10 + ((i * 3 + j) * 3 + k) * 3 + l) * n + qidx] =
                Number(2) * a.lambda * Fi[j][i] * Fi[l][k] + scalar * Fi[l][i] * Fi[j][k] +
                ((i == k && j == l) ? a.mu : Number(0));
    return;
  }
  inv3(Jc, Jci);
  const Number jxw = (Number)wq * fabs(det3(Jc));
  // tensor2, formulation 1 keeps its first 16 fields in pairs (see QSolidT2F1)
  const bool paired = a.formulation == 1 && a.caching == MFG_CACHING_TENSOR2;
  auto at = [&](int f) -> int64_t { return paired && f < 16 ? ((f >> 1) * n + qidx) * 2 + (f & 1) : f * n + qidx; };
#pragma unroll
  for (int f = 0; f < 9; ++f) a.qd[at(f)] = Jci[f / 3][f % 3];
  if (a.formulation == 0) {
    // mf_nh_operator.h:419-440: J^{-1/3} and F
    a.qd[9 * n + qidx] = jxw;
    a.qd[10 * n + qidx] = pow(J, Number(-1.0 / 3.0));
#pragma unroll
    for (int f = 0; f < 9; ++f) a.qd[(11 + f) * n + qidx] = F[f / 3][f % 3];
    return;
  }
  const Number invJ = Number(1) / J;
  Number b[3][3];
#pragma unroll
  for (int x = 0; x < 3; ++x)
#pragma unroll
    for (int y = 0; y < 3; ++y) b[x][y] = F[x][0] * F[y][0] + F[x][1] * F[y][1] + F[x][2] * F[y][2];
  if (a.caching == MFG_CACHING_TENSOR4) {  // mf_nh_operator.h:478-494
    a.qd[9 * n + qidx] = jxw;
    a.qd[10 * n + qidx] = invJ;
#pragma unroll
    for (int s6 = 0; s6 < 6; ++s6) {
      const Number tau = a.mu * b[kSymI[s6]][kSymJ[s6]] - (kSymI[s6] == kSymJ[s6] ? scalar : Number(0));
      a.qd[(11 + s6) * n + qidx] = tau / J;
    }
    // C_ijkl = (2 s_c / J) (d_ik d_jl + d_il d_jk)/2 + (2 lambda / J) d_ij d_kl over symmetric pairs
#pragma unroll
    for (int s6 = 0; s6 < 6; ++s6)
#pragma unroll
      for (int t6 = 0; t6 < 6; ++t6) {
        const int i = kSymI[s6], j = kSymJ[s6], k = kSymI[t6], l = kSymJ[t6];
        const Number S = Number(0.5) * (Number((i == k && j == l) ? 1 : 0) + Number((i == l && j == k) ? 1 : 0));
        const Number IxI = (i == j && k == l) ? Number(1) : Number(0);
        a.qd[(17 + 6 * s6 + t6) * n + qidx] = (scalar * Number(2) / J) * S + (a.lambda * Number(2) / J) * IxI;
      }
    return;
  }
  // tensor2, formulation 1 (mf_nh_operator.h:456-477): the reference's
  // 2 s / J and tau / J times the current JxW = JxW_ref J, i.e. JxW_ref s and
  // JxW_ref tau, plus JxW_ref itself unless the operator takes it per cell
  const double detr = Jr[0][0] * (Jr[1][1] * Jr[2][2] - Jr[1][2] * Jr[2][1]) -
                      Jr[0][1] * (Jr[1][0] * Jr[2][2] - Jr[1][2] * Jr[2][0]) +
                      Jr[0][2] * (Jr[1][0] * Jr[2][1] - Jr[1][1] * Jr[2][0]);
  const Number jxw_ref = (Number)(wq * fabs(detr));
  a.qd[at(9)] = jxw_ref * scalar;
#pragma unroll
  for (int s6 = 0; s6 < 6; ++s6) {
    const Number tau = a.mu * b[kSymI[s6]][kSymJ[s6]] - (kSymI[s6] == kSymJ[s6] ? scalar : Number(0));
    a.qd[at(10 + s6)] = jxw_ref * tau;
  }
  if (a.jxw_ref_field) a.qd[at(16)] = jxw_ref;
}

// ---------------------------------------------------------- diagonal ---
// The diagonal entry of dof (node a, component d) is
//   sum_q [ m_d(q) phi_a(q)^2 + sum_{f,g} G_d[f][g](q) dphi_a/dxi_f dphi_a/dxi_g ],
// where m_d and G_d come from applying the quadrature op to unit values and
// unit reference gradients.  Each of the 7 (m, G_ff, G_fg+G_gf) terms is a
// tensor-product sum over q with squared 1D matrices (phi^2, phi*phi',
// phi'^2), evaluated with the same transposed sweeps as integrate.
template <typename Number>
struct DiagArgs {
  const uint32_t *cell_nodes;
  int cell_begin, cell_end;  // one colour: cells share no node
  const Number *sq;  // 3 x N x N: SS, SD, DD
  const Number *shape, *colloc;  // for ops that read a nodal field
  Number *diag;
};

template <int P, typename Number, class QOp>
__global__ void __launch_bounds__(((P + 1) * (P + 1) * (P + 1) + 63) / 64 * 64) diagonal_cells(DiagArgs<Number> a, QOp qop) {
  constexpr int N = P + 1, N3 = N * N * N, NC = QOp::NC, NTERM = 7 * NC;
  __shared__ Number s_sq[3 * N * N];
  __shared__ Number s_tile[NTERM * N3];
  const int t = threadIdx.x;
  for (int x = t; x < 3 * N * N; x += blockDim.x) s_sq[x] = a.sq[x];
  const bool pad = t >= N3;
  const int pt = pad ? 0 : t;
  Tile<N, Number> T;
  T.base = s_tile;
  T.pt = pt;
  T.i = pt % N;
  T.j = (pt / N) % N;
  T.k = pt / (N * N);
  T.writer = !pad;
  const int cell = a.cell_begin + blockIdx.x;
  const bool active = !pad && cell < a.cell_end;
  const int64_t qidx = active ? (int64_t)cell * N3 + pt : 0;
  Number coef[NTERM];
  {
    typename QOp::Data qd = qop.load(qidx, active ? cell : 0, pt);
    if constexpr (QOp::kNodalField) {
      __shared__ Number s_shape[N * N], s_colloc[N * N];
      if (t < N * N) {
        s_shape[t] = a.shape[t];
        s_colloc[t] = a.colloc[t];
      }
      const int64_t node = active ? (int64_t)(a.cell_nodes[qidx] & kNodeMaskDev) : 0;
      Number w[3], wg[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) w[c] = active ? qop.u[node * 3 + c] : Number(0);
      evaluate_g<N, 3, Number>(T, s_shape, s_colloc, w, wg);
      qop.attach(qd, wg);
    }
    Number v[NC], g[NC][3];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      v[c] = 1;
      g[c][0] = g[c][1] = g[c][2] = 0;
    }
    qop.apply(qd, v, g);
#pragma unroll
    for (int c = 0; c < NC; ++c) coef[c * 7] = v[c];
    Number G[NC][3][3];
#pragma unroll
    for (int d = 0; d < NC; ++d)
#pragma unroll
      for (int gg = 0; gg < 3; ++gg) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          v[c] = 0;
          g[c][0] = g[c][1] = g[c][2] = 0;
        }
        g[d][gg] = 1;
        qop.apply(qd, v, g);
#pragma unroll
        for (int f = 0; f < 3; ++f) G[d][f][gg] = g[d][f];
      }
#pragma unroll
    for (int d = 0; d < NC; ++d) {
      coef[d * 7 + 1] = G[d][0][0];
      coef[d * 7 + 2] = G[d][1][1];
      coef[d * 7 + 3] = G[d][2][2];
      coef[d * 7 + 4] = G[d][0][1] + G[d][1][0];
      coef[d * 7 + 5] = G[d][0][2] + G[d][2][0];
      coef[d * 7 + 6] = G[d][1][2] + G[d][2][1];
    }
    if (!active)
#pragma unroll
      for (int x = 0; x < NTERM; ++x) coef[x] = 0;
  }
  // matrix (0 SS, 1 SD, 2 DD) per term and direction
  constexpr int mat[7][3] = {{0, 0, 0}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
  __syncthreads();
#pragma unroll
  for (int x = 0; x < NTERM; ++x) T.put(x, coef[x]);
  __syncthreads();
#pragma unroll
  for (int x = 0; x < NTERM; ++x) coef[x] = T.template line_t<2>(s_sq + mat[x % 7][2] * N * N, T.k, x);
  __syncthreads();
#pragma unroll
  for (int x = 0; x < NTERM; ++x) T.put(x, coef[x]);
  __syncthreads();
#pragma unroll
  for (int x = 0; x < NTERM; ++x) coef[x] = T.template line_t<1>(s_sq + mat[x % 7][1] * N * N, T.j, x);
  __syncthreads();
#pragma unroll
  for (int x = 0; x < NTERM; ++x) T.put(x, coef[x]);
  __syncthreads();
  if (!active) return;
  const uint32_t entry = a.cell_nodes[qidx];
  const int64_t node = entry & kNodeMaskDev;
#pragma unroll
  for (int d = 0; d < NC; ++d) {
    Number s = 0;
#pragma unroll
    for (int x = 0; x < 7; ++x) s += T.template line_t<0>(s_sq + mat[x][0] * N * N, T.i, d * 7 + x);
    if (!((entry >> (kConstraintShift + d)) & 1u)) a.diag[node * NC + d] += s;
  }
}

}  // namespace mfg
