// CPU restatement of the reference's tangent vmult with the reference's own
// CPU strategy, for the CPU baseline of the benchmark: cells in batches of 8
// processed together in SIMD lanes (deal.II's VectorizedArray<double> with
// AVX-512), sum factorisation with the collocation derivative at the Gauss
// points (FEEvaluation's evaluate/integrate for n_q = p+1), OpenMP threads
// over the batches of one colour (cells of a colour share no node, so the
// scatter needs no atomics).  Tensor2 caching, formulation 1
// (mf_nh_operator.h:945-974, the point data cached at :456-477); the
// constrained rows are the identity (:622-624).
//
// Test infrastructure like the rest of this directory: only the tests and the
// benchmark's CPU baseline call it.  It takes the oracle's 18 canonical words
// per point (orc_solid_tensor2_data) and the mesh arrays, so it shares no code
// with the device library.  [pinned: equals orc_solid_vmult to 1e-13 relative,
// tests/test_oracle_operators.py]
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "fe.hpp"

namespace {

constexpr int W = 8;  // lanes: cells per batch
typedef double v8d __attribute__((vector_size(W * sizeof(double))));

inline v8d splat(double x) { return v8d{x, x, x, x, x, x, x, x}; }

struct SimdSolid {
  int p = 1, N = 2, N3 = 8;
  int n[3] = {0, 0, 0};
  int64_t n_cells = 0, n_dofs = 0;
  std::vector<int32_t> cell_nodes;  // lexicographic per cell, N3 entries
  std::vector<uint8_t> con;         // per dof
  std::vector<int32_t> constrained;
  double two_lambda = 0, rho_alpha = 0;
  // batches: cells (-1 pads the last batch of a colour), colour ranges
  std::vector<int32_t> batch_cells;
  std::vector<int> colour_begin;
  // point data per batch: [batch][q][18] lanes
  std::vector<v8d> qd;
  // 1D tables: S[q][a] (values of the GLL basis at the Gauss points) and
  // D[q][r] (derivative of the Lagrange basis on the Gauss points at point q)
  std::vector<double> S, D;
};

// out[.. i ..] = sum_a M[i][a] in[.. a ..] along direction DIR of an N^3 tile
template <int NN, int DIR, bool TRANSPOSE>
inline void sweep(const double *M, const v8d *in, v8d *out) {
  constexpr int S0 = DIR == 0 ? 1 : DIR == 1 ? NN : NN * NN;
  constexpr int O1 = DIR == 0 ? NN : 1, O2 = DIR == 2 ? NN : NN * NN;
  for (int b = 0; b < NN; ++b)
    for (int a = 0; a < NN; ++a) {
      const int base = a * O1 + b * O2;
      v8d x[NN];
#pragma GCC unroll 8
      for (int k = 0; k < NN; ++k) x[k] = in[base + k * S0];
#pragma GCC unroll 8
      for (int i = 0; i < NN; ++i) {
        v8d s = splat(0.0);
#pragma GCC unroll 8
        for (int k = 0; k < NN; ++k) s += splat(TRANSPOSE ? M[k * NN + i] : M[i * NN + k]) * x[k];
        out[base + i * S0] = s;
      }
    }
}

template <int NN>
void apply_batch(const SimdSolid &op, int batch, const double *src, double *dst) {
  constexpr int M3 = NN * NN * NN;
  const int32_t *cells = &op.batch_cells[(size_t)batch * W];
  alignas(64) v8d u[3][M3], t[M3], uq[3][M3], g[3][3][M3], acc[M3];
  // gather, constrained entries read as zero
  for (int i = 0; i < M3; ++i)
    for (int d = 0; d < 3; ++d) {
      v8d v = splat(0.0);
      for (int l = 0; l < W; ++l) {
        const int c = cells[l];
        if (c < 0) continue;
        const int64_t dof = 3 * (int64_t)op.cell_nodes[(size_t)c * M3 + i] + d;
        v[l] = op.con[dof] ? 0.0 : src[dof];
      }
      u[d][i] = v;
    }
  // evaluate: values at the Gauss points, then the collocation derivative
  const double *S = op.S.data(), *D = op.D.data();
  for (int d = 0; d < 3; ++d) {
    sweep<NN, 0, false>(S, u[d], t);
    sweep<NN, 1, false>(S, t, acc);
    sweep<NN, 2, false>(S, acc, uq[d]);
    sweep<NN, 0, false>(D, uq[d], g[d][0]);
    sweep<NN, 1, false>(D, uq[d], g[d][1]);
    sweep<NN, 2, false>(D, uq[d], g[d][2]);
  }
  // point operation (mf_nh_operator.h:945-974)
  const v8d *Q = &op.qd[(size_t)batch * M3 * 18];
  const v8d tl = splat(op.two_lambda), ra = splat(op.rho_alpha);
  for (int q = 0; q < M3; ++q) {
    const v8d *f = Q + (size_t)q * 18;
    v8d gx[3][3];
    for (int c = 0; c < 3; ++c)
      for (int e = 0; e < 3; ++e) gx[c][e] = g[c][0][q] * f[0 * 3 + e] + g[c][1][q] * f[1 * 3 + e] + g[c][2][q] * f[2 * 3 + e];
    const v8d jxw = f[9], c2 = f[10], c1 = f[11];
    const v8d T[3][3] = {{f[12], f[15], f[16]}, {f[15], f[13], f[17]}, {f[16], f[17], f[14]}};
    const v8d tr = gx[0][0] + gx[1][1] + gx[2][2];
    v8d P[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        v8d x = c2 * (splat(0.5) * (gx[a][b] + gx[b][a])) + gx[a][0] * T[0][b] + gx[a][1] * T[1][b] + gx[a][2] * T[2][b];
        if (a == b) x += tl * c1 * tr;
        P[a][b] = x * jxw;
      }
    for (int c = 0; c < 3; ++c) {
      for (int e = 0; e < 3; ++e) g[c][e][q] = P[c][0] * f[e * 3 + 0] + P[c][1] * f[e * 3 + 1] + P[c][2] * f[e * 3 + 2];
      uq[c][q] = uq[c][q] * ra * c1 * jxw;
    }
  }
  // integrate: transposed collocation derivative, then the transposed values
  for (int d = 0; d < 3; ++d) {
    sweep<NN, 0, true>(D, g[d][0], t);
    for (int i = 0; i < M3; ++i) uq[d][i] += t[i];
    sweep<NN, 1, true>(D, g[d][1], t);
    for (int i = 0; i < M3; ++i) uq[d][i] += t[i];
    sweep<NN, 2, true>(D, g[d][2], t);
    for (int i = 0; i < M3; ++i) uq[d][i] += t[i];
    sweep<NN, 2, true>(S, uq[d], t);
    sweep<NN, 1, true>(S, t, acc);
    sweep<NN, 0, true>(S, acc, u[d]);
  }
  // scatter: no other batch of this colour touches these nodes
  for (int l = 0; l < W; ++l) {
    const int c = cells[l];
    if (c < 0) continue;
    for (int i = 0; i < M3; ++i)
      for (int d = 0; d < 3; ++d) {
        const int64_t dof = 3 * (int64_t)op.cell_nodes[(size_t)c * M3 + i] + d;
        if (!op.con[dof]) dst[dof] += u[d][i][l];
      }
  }
}

template <int NN>
void vmult_t(const SimdSolid &op, double *dst, const double *src) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < op.n_dofs; ++i) dst[i] = 0.0;
  for (size_t col = 0; col + 1 < op.colour_begin.size(); ++col) {
#pragma omp parallel for schedule(static)
    for (int b = op.colour_begin[col]; b < op.colour_begin[col + 1]; ++b) apply_batch<NN>(op, b, src, dst);
  }
  for (int32_t c : op.constrained) dst[c] += src[c];
}

}  // namespace

extern "C" {

// Mesh: n[3] cells per direction (cell = cx + n0 (cy + n1 cz)), degree p with
// n_q = p + 1, cell_nodes (N^3 per cell, lexicographic), the sorted
// constrained dof list.  fields: the 18 canonical words per point, field-major
// [f][cell * N^3 + q] (Jinv_cur 9, JxW_cur, 2 s / J, 1 / J, tau / J as xx yy
// zz xy xz yz).  Degrees 2 and 4 (the baseline's configurations); other
// degrees return NULL.
void *orc_simd_solid_create(int p, const int *n, int64_t n_dofs, const int32_t *cell_nodes, int64_t n_constrained,
                            const int32_t *constrained, const double *fields, double mu, double nu, double rho_alpha) {
  if (p != 2 && p != 4) return nullptr;
  auto *op = new SimdSolid();
  op->p = p;
  op->N = p + 1;
  op->N3 = op->N * op->N * op->N;
  for (int d = 0; d < 3; ++d) op->n[d] = n[d];
  op->n_cells = (int64_t)n[0] * n[1] * n[2];
  op->n_dofs = n_dofs;
  op->cell_nodes.assign(cell_nodes, cell_nodes + op->n_cells * op->N3);
  op->con.assign((size_t)n_dofs, 0);
  op->constrained.assign(constrained, constrained + n_constrained);
  for (int32_t c : op->constrained) op->con[c] = 1;
  op->two_lambda = 2.0 * (2.0 * mu * nu / (1.0 - 2.0 * nu));
  op->rho_alpha = rho_alpha;
  // 1D tables on Gauss points
  const int N = op->N;
  std::vector<double> gll(N), qx(N), qw(N);
  orc::gll_points(p, gll.data());
  orc::gauss_rule(N, qx.data(), qw.data());
  op->S.resize(N * N);
  op->D.resize(N * N);
  std::vector<double> der(N);
  for (int q = 0; q < N; ++q) {
    orc::lagrange_eval(N, gll.data(), qx[q], &op->S[q * N], der.data());
    std::vector<double> val(N);
    orc::lagrange_eval(N, qx.data(), qx[q], val.data(), &op->D[q * N]);
  }
  // batches of 8 cells of one parity colour
  op->colour_begin.assign(1, 0);
  for (int col = 0; col < 8; ++col) {
    std::vector<int32_t> cs;
    for (int64_t c = 0; c < op->n_cells; ++c) {
      const int cx = (int)(c % n[0]), cy = (int)((c / n[0]) % n[1]), cz = (int)(c / ((int64_t)n[0] * n[1]));
      if (((cx & 1) | ((cy & 1) << 1) | ((cz & 1) << 2)) == col) cs.push_back((int32_t)c);
    }
    while (cs.size() % W) cs.push_back(-1);
    op->batch_cells.insert(op->batch_cells.end(), cs.begin(), cs.end());
    op->colour_begin.push_back((int)(op->batch_cells.size() / W));
  }
  const int nb = op->colour_begin.back();
  const int64_t nqp = op->n_cells * op->N3;
  op->qd.assign((size_t)nb * op->N3 * 18, splat(0.0));
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b)
    for (int l = 0; l < W; ++l) {
      const int c = op->batch_cells[(size_t)b * W + l];
      for (int q = 0; q < op->N3; ++q)
        for (int f = 0; f < 18; ++f)
          op->qd[((size_t)b * op->N3 + q) * 18 + f][l] = c < 0 ? 0.0 : fields[(size_t)f * nqp + (size_t)c * op->N3 + q];
    }
  return op;
}

void orc_simd_solid_destroy(void *op) { delete (SimdSolid *)op; }

void orc_simd_solid_vmult(void *h, double *dst, const double *src) {
  const SimdSolid &op = *(const SimdSolid *)h;
  if (op.p == 2) vmult_t<3>(op, dst, src);
  else vmult_t<5>(op, dst, src);
}

int orc_simd_lanes(void) { return W; }

}  // extern "C"
