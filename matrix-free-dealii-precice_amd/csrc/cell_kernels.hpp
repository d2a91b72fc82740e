// Sum-factorisation cell kernels for gfx950 (CDNA4, wave64).
//
// Layout of one launch: a workgroup processes CPB cells, one thread per
// quadrature point (n_q = p+1, so points and nodes share the N^3 index space
// and the thread also owns node (i,j,k) of its cell).  The three 1D sweeps of
// each evaluate/integrate pass go through an LDS tile per cell; the
// collocation derivative at the Gauss points turns the three gradient
// directions into one LDS round trip.  Gather is a coalesced read of the
// int32 cell->node map followed by per-node vector loads.  Cells come in
// colours (no two cells of a colour share a node), one launch per colour, so
// the scatter is a plain read-modify-write, whose read is issued with the
// gather; the first colour that touches a node stores instead (bit 31 of its
// map entry), which makes vmult need no zero fill.  Constrained dofs are skipped via the constraint bits stored
// above bit 28 of every map entry.
//
// Algorithmic reference: FEEvaluation::evaluate/integrate with
// values|gradients as used by NeoHookOperator::do_operation_on_cell
// (mf_nh_operator.h:651-1019) and LaplaceOperator::do_operation_on_cell
// (heat_transfer.h:234-250).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace mfg {

constexpr int kConstraintShift = 28;
constexpr uint32_t kNodeMaskDev = (1u << kConstraintShift) - 1u;

template <int P, int CPB>
constexpr int block_threads() {
  return ((CPB * (P + 1) * (P + 1) * (P + 1) + 63) / 64) * 64;
}

// cells per workgroup of the apply / cache kernels, per degree.  Q4: one cell
// per workgroup measured 3-5 % faster than two for the solid and the heat
// operator in FP64 and, once the per-cell tables became scalar loads, 2-5 %
// in FP32.  Q3 and Q5 with one cell instead of two, Q2 with two cells (one
// wave) instead of seven, on 3-6M-DoF meshes: solid FP64 -4 to -30 %, heat
// -4 to -26 %, FP32 -1 to -3 %, small meshes unchanged.
template <int P, typename Number>
struct DefaultCPB {
  static constexpr int value = P == 1 ? 8 : P == 2 ? 2 : 1;
};

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each
// XCD walks one contiguous range of cells (shared nodes stay in one L2).
__device__ inline int xcd_remap(int b, int G) {
  const int x = b & 7, l = b >> 3, q = G >> 3, r = G & 7;
  return x * q + (x < r ? x : r) + l;
}


// Per-thread view of the cell tile in LDS.
template <int N, typename Number>
struct Tile {
  static constexpr int N3 = N * N * N;
  Number *base;  // slot 0 of this cell
  int i, j, k, pt;
  bool writer;   // false for padding lanes
  __device__ Number *slot(int s) const { return base + s * N3; }
  __device__ void put(int s, Number v) const {
    if (writer) base[s * N3 + pt] = v;
  }
  // sum_a M[row*N + a] * slot(a along dir)
  template <int DIR>
  __device__ Number line(const Number *M, int row, int s) const {
    const Number *src = slot(s);
    Number r = 0;
    if (DIR == 0) {
      const Number *b = src + N * (j + N * k);
#pragma unroll
      for (int a = 0; a < N; ++a) r += M[row * N + a] * b[a];
    } else if (DIR == 1) {
      const Number *b = src + i + N * N * k;
#pragma unroll
      for (int a = 0; a < N; ++a) r += M[row * N + a] * b[a * N];
    } else {
      const Number *b = src + i + N * j;
#pragma unroll
      for (int a = 0; a < N; ++a) r += M[row * N + a] * b[a * N * N];
    }
    return r;
  }
  // sum_a M[a*N + col] * slot(a along dir)   (transposed matrix)
  template <int DIR>
  __device__ Number line_t(const Number *M, int col, int s) const {
    const Number *src = slot(s);
    Number r = 0;
    if (DIR == 0) {
      const Number *b = src + N * (j + N * k);
#pragma unroll
      for (int a = 0; a < N; ++a) r += M[a * N + col] * b[a];
    } else if (DIR == 1) {
      const Number *b = src + i + N * N * k;
#pragma unroll
      for (int a = 0; a < N; ++a) r += M[a * N + col] * b[a * N];
    } else {
      const Number *b = src + i + N * j;
#pragma unroll
      for (int a = 0; a < N; ++a) r += M[a * N + col] * b[a * N * N];
    }
    return r;
  }
};

// The gradient phase of evaluate_vg and the first phase of integrate_vg read
// 3 NC lines of N LDS values each.  Unrolled over the components, the
// compiler issues all of them at once: for N >= 5 that peak (about 120 VGPRs
// in FP64) holds the cell kernel at 2 waves per SIMD.  A rolled component
// loop frees those registers (3 to 4 waves); for N <= 4 the unrolled form is
// faster.
template <int N>
constexpr bool rolled_components() {
  return N >= 5;
}

// Evaluate values and reference-coordinate gradients at the quadrature points
// of NC components; slots 0..2NC-1 are used.  On return the B slots hold the
// point values (still read by later lines until the next barrier).
template <int N, int NC, typename Number>
__device__ inline void evaluate_vg(const Tile<N, Number> &T, const Number *S, const Number *D, const Number (&u)[NC],
                                   Number (&val)[NC], Number (&grad)[NC][3]) {
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
#pragma unroll
  for (int c = 0; c < NC; ++c) T.put(NC + c, val[c]);
  __syncthreads();
  auto gradient = [&](int c) {
    grad[c][0] = T.template line<0>(D, T.i, NC + c);
    grad[c][1] = T.template line<1>(D, T.j, NC + c);
    grad[c][2] = T.template line<2>(D, T.k, NC + c);
  };
  if constexpr (rolled_components<N>()) {
#pragma unroll 1
    for (int c = 0; c < NC; ++c) gradient(c);
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c) gradient(c);
  }
}

// Integrate submitted values and reference gradients back to the nodes
// (transpose of evaluate_vg).  Uses all 3NC slots; with kSpareSlots the tile
// has 5NC slots and the gradients go to slots 2NC.. that evaluate_vg does not
// use, which makes the two barriers that only guard slot reuse (after
// evaluate_vg's last reads and after the gradient lines) unnecessary.
template <int N, int NC, typename Number, bool kSpareSlots = false>
__device__ inline void integrate_vg(const Tile<N, Number> &T, const Number *S, const Number *D,
                                    const Number (&val)[NC], const Number (&grad)[NC][3], Number (&out)[NC]) {
  constexpr int G0 = kSpareSlots ? 2 * NC : 0;  // first gradient slot
  if constexpr (!kSpareSlots) __syncthreads();  // B slots may still be read by evaluate_vg's last lines
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int f = 0; f < 3; ++f) T.put(G0 + 3 * c + f, grad[c][f]);
  __syncthreads();
  Number r[NC];
  auto gradient_t = [&](int c) {
    r[c] = val[c] + T.template line_t<0>(D, T.i, G0 + 3 * c + 0) + T.template line_t<1>(D, T.j, G0 + 3 * c + 1) +
           T.template line_t<2>(D, T.k, G0 + 3 * c + 2);
  };
  if constexpr (rolled_components<N>()) {
#pragma unroll 1
    for (int c = 0; c < NC; ++c) gradient_t(c);
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c) gradient_t(c);
  }
  if constexpr (!kSpareSlots) __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; ++c) T.put(c, r[c]);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; ++c) r[c] = T.template line_t<2>(S, T.k, c);
#pragma unroll
  for (int c = 0; c < NC; ++c) T.put(NC + c, r[c]);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; ++c) r[c] = T.template line_t<1>(S, T.j, NC + c);
#pragma unroll
  for (int c = 0; c < NC; ++c) T.put(c, r[c]);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; ++c) out[c] = T.template line_t<0>(S, T.i, c);
}

// Evaluate only reference gradients (cache()); same slot usage as evaluate_vg.
template <int N, int NC, typename Number>
__device__ inline void evaluate_g(const Tile<N, Number> &T, const Number *S, const Number *D, const Number (&u)[NC],
                                  Number (&grad)[NC][3]) {
  Number val[NC];
  evaluate_vg<N, NC, Number>(T, S, D, u, val, grad);
}

// The quadrature data is read once per operator application: load it
// non-temporally so that it does not evict from the caches the vector entries
// that the next colour reads again (nodes shared between cells).
template <typename Number>
__device__ inline Number stream_load(const Number *p) {
  return __builtin_nontemporal_load(p);
}

// Per-cell tables (geometry of affine cells) are read-only during a launch.
// Read through the constant address space: where the cell index is uniform
// across the wave (one cell per workgroup) the loads become scalar loads into
// SGPRs, otherwise they stay vector loads.
template <typename Number>
using ConstantPtr = const Number __attribute__((address_space(4))) *;
template <typename Number>
__device__ inline ConstantPtr<Number> as_constant(const Number *p) {
  return (ConstantPtr<Number>)p;
}

// ------------------------------------------------------ quadrature ops ---
// Each op provides load(qidx, cell, pt) -> Data (point qidx = cell N^3 + pt
// of the stored cell order) and apply(Data, val, grad), turning values and
// reference gradients at a point into submitted value and reference gradient
// already multiplied by JxW.

// symmetric index pairs xx, yy, zz, xy, xz, yz
constexpr int kSymI[6] = {0, 1, 2, 0, 0, 1}, kSymJ[6] = {0, 1, 2, 1, 2, 2};

// Solid, tensor2 caching, formulation 1 (mf_nh_operator.h:945-974).
// The reference caches, per point, the current-frame Jinv and JxW (from
// MappingQEulerian) and the scalars 2 s / J, 1 / J and tau / J, where
// s = mu - 2 lambda ln J and tau is the Kirchhoff stress.  The operator only
// uses them multiplied by the current JxW = JxW_ref J, so these products are
// cached instead:
//   fields 0..8   Jinv_cur[f][e]
//   field  9      JxW_ref s
//   fields 10..15 JxW_ref tau (xx, yy, zz, xy, xz, yz)
//   field  16     JxW_ref, only on meshes with a cell that is not affine in the
//                 reference configuration.  Otherwise JxW_ref = w_q det J_ref
//                 comes from one word per cell (cell_jxw) and the tensor
//                 weights (wq).
// 16 words per point instead of the reference's 18, and no JxW products in
// the point operation.
template <typename Number>
struct QSolidT2F1 {
  static constexpr int NC = 3;
  static constexpr int kFields = 17;  // 16 with cell_jxw
  static constexpr bool kNodalField = false;
  const Number *qd;
  int64_t nqp;
  Number two_lambda, rho_alpha;
  const Number *cell_jxw;  // det J_ref per stored cell, or null: field 16 holds JxW_ref
  const Number *wq;        // N^3 tensor-product quadrature weights (with cell_jxw)
  struct Data {
    Number ji[9], s, T[6], w;
  };
  // fields 0..15 are stored in pairs, qd[((f / 2) nqp + q) 2 + f % 2], field
  // 16 plainly after them: 8 16-byte (FP64) loads bring a point's 16 words
  typedef Number Pair __attribute__((ext_vector_type(2)));
  __device__ Data load(int64_t q, int cell, int pt) const {
    Data d;
    Number f16[16];
    const Pair *qp = reinterpret_cast<const Pair *>(qd);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const Pair x = __builtin_nontemporal_load(qp + k * nqp + q);
      f16[2 * k] = x.x;
      f16[2 * k + 1] = x.y;
    }
#pragma unroll
    for (int f = 0; f < 9; ++f) d.ji[f] = f16[f];
    d.s = f16[9];
#pragma unroll
    for (int f = 0; f < 6; ++f) d.T[f] = f16[10 + f];
    d.w = cell_jxw ? as_constant(cell_jxw)[cell] * wq[pt] : stream_load(qd + 16 * nqp + q);
    return d;
  }
  __device__ void apply(const Data &d, Number (&val)[3], Number (&grad)[3][3]) const {
    Number g[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < 3; ++e) g[c][e] = grad[c][0] * d.ji[0 * 3 + e] + grad[c][1] * d.ji[1 * 3 + e] + grad[c][2] * d.ji[2 * 3 + e];
    const Number tr = g[0][0] + g[1][1] + g[2][2];
    const Number diag = two_lambda * d.w * tr;
    // T symmetric: T[0]=xx T[1]=yy T[2]=zz T[3]=xy T[4]=xz T[5]=yz
    const Number Tm[3][3] = {{d.T[0], d.T[3], d.T[4]}, {d.T[3], d.T[1], d.T[5]}, {d.T[4], d.T[5], d.T[2]}};
    Number Pm[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        Pm[a][b] = d.s * (g[a][b] + g[b][a]) + (a == b ? diag : Number(0)) + g[a][0] * Tm[0][b] +
                   g[a][1] * Tm[1][b] + g[a][2] * Tm[2][b];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int f = 0; f < 3; ++f)
        grad[c][f] = Pm[c][0] * d.ji[f * 3 + 0] + Pm[c][1] * d.ji[f * 3 + 1] + Pm[c][2] * d.ji[f * 3 + 2];
      val[c] = rho_alpha * d.w * val[c];
    }
  }
};

// Solid, tensor2 caching, formulation 1, rebuilt per point from the
// displacement gradient: the operator of QSolidT2F1 from 10 words per point.
// On affine meshes J_ref and J_ref^{-1} are constant per cell, so what the
// folded layout stores follows from Grad_xi u (the reference-cell gradient of
// the linearisation point, 9 words) and s = mu - 2 lambda ln J (1 word) with
// the arithmetic of the cache kernel:
//   J_cur = J_ref + Grad_xi u,   Jinv_cur = J_cur^{-1},
//   F = I + Grad_xi u J_ref^{-1},   tau = mu F F^T - s I,
// and JxW_ref = det J_ref w_q.  Fields: Grad_xi u [d][f] (9), s; stored in
// pairs of fields (below).
// kEarlyCell (the cell index is the same on every lane: one cell per
// workgroup): the cell's J_ref, J_ref^{-1} and det J_ref are loaded into
// SGPRs with the point data, so that their latency overlaps the gather's
// instead of following the evaluate sweeps.
// kCartesian (with kEarlyCell; every cell's J_ref is diagonal, as on
// axis-aligned boxes): only the 3 + 3 diagonal entries are loaded, and the
// products with the exact zeros off the diagonal are left out.  Every
// rounding step of the general form survives, so the two agree bit for bit
// (up to the sign of a zero).
template <typename Number, bool kEarlyCell = false, bool kCartesian = false>
struct QSolidT2F1G {
  static_assert(kEarlyCell || !kCartesian, "the Cartesian form loads the cell's diagonal into SGPRs");
  static constexpr int NC = 3;
  static constexpr int kFields = 10;
  static constexpr bool kNodalField = false;
  const Number *qd;
  int64_t nqp;
  Number mu, two_lambda, rho_alpha;
  const Number *cell_jac;  // per stored cell: J_ref (9), J_ref^{-1} (9)
  const Number *cell_jxw;  // det J_ref per stored cell
  const Number *wq;        // N^3 tensor-product quadrature weights
  struct Data {
    Number gu[9], s, w;
    const Number *jac;
    // kCartesian: diag J_ref (3), diag J_ref^{-1} (3); else J_ref (9), J_ref^{-1} (9)
    Number jr[kEarlyCell ? (kCartesian ? 6 : 18) : 1];
  };
  // fields are stored in pairs, qd[((f / 2) nqp + q) 2 + f % 2]: one 16-byte
  // (FP64) load brings two fields of a point
  typedef Number Pair __attribute__((ext_vector_type(2)));
  __device__ Data load(int64_t q, int cell, int pt) const {
    Data d;
    const Pair *qp = reinterpret_cast<const Pair *>(qd);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const Pair x = __builtin_nontemporal_load(qp + k * nqp + q);
      if (k < 4) {
        d.gu[2 * k] = x.x;
        d.gu[2 * k + 1] = x.y;
      } else {
        d.gu[8] = x.x;
        d.s = x.y;
      }
    }
    load_cell(d, cell, pt);
    return d;
  }
  // the per-cell part of load(): J_ref, J_ref^{-1} and JxW_ref at point pt
  __device__ void load_cell(Data &d, int cell, int pt) const {
    d.jac = cell_jac + (int64_t)cell * 18;
    if constexpr (kCartesian) {
      Number jxw = as_constant(cell_jxw)[cell];
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        d.jr[e] = as_constant(d.jac)[4 * e];
        d.jr[3 + e] = as_constant(d.jac)[9 + 4 * e];
      }
      asm volatile("" : "+s"(jxw), "+s"(d.jr[0]), "+s"(d.jr[1]), "+s"(d.jr[2]), "+s"(d.jr[3]), "+s"(d.jr[4]),
                   "+s"(d.jr[5]));
      d.w = jxw * wq[pt];
    } else if constexpr (kEarlyCell) {
      Number jxw = as_constant(cell_jxw)[cell];
#pragma unroll
      for (int f = 0; f < 18; ++f) d.jr[f] = as_constant(d.jac)[f];
      // all 19 values in SGPRs here: one wait, before the first barrier
      asm volatile(""
                   : "+s"(jxw), "+s"(d.jr[0]), "+s"(d.jr[1]), "+s"(d.jr[2]), "+s"(d.jr[3]), "+s"(d.jr[4]),
                     "+s"(d.jr[5]), "+s"(d.jr[6]), "+s"(d.jr[7]), "+s"(d.jr[8]), "+s"(d.jr[9]), "+s"(d.jr[10]),
                     "+s"(d.jr[11]), "+s"(d.jr[12]), "+s"(d.jr[13]), "+s"(d.jr[14]), "+s"(d.jr[15]), "+s"(d.jr[16]),
                     "+s"(d.jr[17]));
      d.w = jxw * wq[pt];
    } else {
      d.w = as_constant(cell_jxw)[cell] * wq[pt];
    }
  }
  __device__ void apply(const Data &d, Number (&val)[3], Number (&grad)[3][3]) const {
    // J_ref (f < 9) and J_ref^{-1} (9 + f) of the cell
    auto jr = [&](int f) -> Number {
      if constexpr (kEarlyCell) return d.jr[f];
      else return as_constant(d.jac)[f];
    };
    Number ji[9];
    {
      Number Jc[3][3];
      if constexpr (kCartesian) {
#pragma unroll
        for (int f = 0; f < 9; ++f) Jc[f / 3][f % 3] = f % 4 == 0 ? d.jr[f / 4] + d.gu[f] : d.gu[f];
      } else {
#pragma unroll
        for (int f = 0; f < 9; ++f) Jc[f / 3][f % 3] = jr(f) + d.gu[f];
      }
      const Number id =
          Number(1) / (Jc[0][0] * (Jc[1][1] * Jc[2][2] - Jc[1][2] * Jc[2][1]) -
                       Jc[0][1] * (Jc[1][0] * Jc[2][2] - Jc[1][2] * Jc[2][0]) + Jc[0][2] * (Jc[1][0] * Jc[2][1] - Jc[1][1] * Jc[2][0]));
      ji[0] = (Jc[1][1] * Jc[2][2] - Jc[1][2] * Jc[2][1]) * id;
      ji[1] = (Jc[0][2] * Jc[2][1] - Jc[0][1] * Jc[2][2]) * id;
      ji[2] = (Jc[0][1] * Jc[1][2] - Jc[0][2] * Jc[1][1]) * id;
      ji[3] = (Jc[1][2] * Jc[2][0] - Jc[1][0] * Jc[2][2]) * id;
      ji[4] = (Jc[0][0] * Jc[2][2] - Jc[0][2] * Jc[2][0]) * id;
      ji[5] = (Jc[0][2] * Jc[1][0] - Jc[0][0] * Jc[1][2]) * id;
      ji[6] = (Jc[1][0] * Jc[2][1] - Jc[1][1] * Jc[2][0]) * id;
      ji[7] = (Jc[0][1] * Jc[2][0] - Jc[0][0] * Jc[2][1]) * id;
      ji[8] = (Jc[0][0] * Jc[1][1] - Jc[0][1] * Jc[1][0]) * id;
    }
    Number T[6];
    {
      Number F[3][3];
      if constexpr (kCartesian) {
        // the general form's sum of three products, two of them exact zeros,
        // rounds as this one product does
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int e = 0; e < 3; ++e) {
            const Number t = d.gu[a * 3 + e] * d.jr[3 + e];
            F[a][e] = a == e ? t + Number(1) : t;
          }
      } else {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int e = 0; e < 3; ++e)
            F[a][e] = d.gu[a * 3 + 0] * jr(9 + 0 * 3 + e) + d.gu[a * 3 + 1] * jr(9 + 1 * 3 + e) +
                      d.gu[a * 3 + 2] * jr(9 + 2 * 3 + e) + (a == e ? Number(1) : Number(0));
      }
#pragma unroll
      for (int s6 = 0; s6 < 6; ++s6) {
        const int x = kSymI[s6], y = kSymJ[s6];
        const Number b = F[x][0] * F[y][0] + F[x][1] * F[y][1] + F[x][2] * F[y][2];
        T[s6] = d.w * (mu * b - (x == y ? d.s : Number(0)));
      }
    }
    const Number sw = d.w * d.s;
    Number g[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < 3; ++e) g[c][e] = grad[c][0] * ji[0 * 3 + e] + grad[c][1] * ji[1 * 3 + e] + grad[c][2] * ji[2 * 3 + e];
    const Number tr = g[0][0] + g[1][1] + g[2][2];
    const Number diag = two_lambda * d.w * tr;
    const Number Tm[3][3] = {{T[0], T[3], T[4]}, {T[3], T[1], T[5]}, {T[4], T[5], T[2]}};
    Number Pm[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        Pm[a][b] = sw * (g[a][b] + g[b][a]) + (a == b ? diag : Number(0)) + g[a][0] * Tm[0][b] + g[a][1] * Tm[1][b] +
                   g[a][2] * Tm[2][b];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int f = 0; f < 3; ++f) grad[c][f] = Pm[c][0] * ji[f * 3 + 0] + Pm[c][1] * ji[f * 3 + 1] + Pm[c][2] * ji[f * 3 + 2];
      val[c] = rho_alpha * d.w * val[c];
    }
  }
};

// Solid, tensor2 caching, formulation 0 (mf_nh_operator.h:701-840).
// Fields: Jinv_cur (9), JxW, J^{-1/3}, F (9).
template <typename Number>
struct QSolidT2F0 {
  static constexpr int NC = 3;
  static constexpr int kFields = 20;
  static constexpr bool kNodalField = false;
  const Number *qd;
  int64_t nqp;
  Number kappa, c_1, rho_alpha;
  struct Data {
    Number ji[9], jxw, jm13, F[9];
  };
  __device__ Data load(int64_t q, int cell, int pt) const {
    Data d;
#pragma unroll
    for (int f = 0; f < 9; ++f) d.ji[f] = stream_load(qd + f * nqp + q);
    d.jxw = stream_load(qd + 9 * nqp + q);
    d.jm13 = stream_load(qd + 10 * nqp + q);
#pragma unroll
    for (int f = 0; f < 9; ++f) d.F[f] = stream_load(qd + (11 + f) * nqp + q);
    return d;
  }
  __device__ void apply(const Data &d, Number (&val)[3], Number (&grad)[3][3]) const {
    Number g[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < 3; ++e) g[c][e] = grad[c][0] * d.ji[0 * 3 + e] + grad[c][1] * d.ji[1 * 3 + e] + grad[c][2] * d.ji[2 * 3 + e];
    const Number *F = d.F;
    const Number det_F = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                         F[2] * (F[3] * F[7] - F[4] * F[6]);
    Number Fb[9];
#pragma unroll
    for (int a = 0; a < 9; ++a) Fb[a] = F[a] * d.jm13;
    Number tau_bar[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        tau_bar[a][b] = Number(2) * c_1 * (Fb[a * 3 + 0] * Fb[b * 3 + 0] + Fb[a * 3 + 1] * Fb[b * 3 + 1] + Fb[a * 3 + 2] * Fb[b * 3 + 2]);
    const Number tr_tau_bar = tau_bar[0][0] + tau_bar[1][1] + tau_bar[2][2];
    const Number tr_tau_bar_dim = tr_tau_bar / Number(3);
    const Number inv_det = Number(1) / det_F;
    const Number dPsi = (kappa / Number(2)) * (det_F - inv_det);
    const Number dPsi_J = dPsi * det_F;
    const Number d2Psi = (kappa / Number(2)) * (Number(1) + inv_det * inv_det);
    Number s[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) s[a][b] = Number(0.5) * (g[a][b] + g[b][a]);
    const Number tr = s[0][0] + s[1][1] + s[2][2];
    Number tau_iso_src = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) tau_iso_src += (tau_bar[a][b] - (a == b ? tr_tau_bar_dim : Number(0))) * s[a][b];
    const Number vol = det_F * (dPsi + det_F * d2Psi) * tr;
    Number Pm[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const Number tau_iso = tau_bar[a][b] - (a == b ? tr_tau_bar_dim : Number(0));
        const Number dev = s[a][b] - (a == b ? tr / Number(3) : Number(0));
        Number jc = s[a][b] * (-dPsi_J * Number(2)) + Number(2.0 / 3.0) * tr_tau_bar * dev -
                    Number(2.0 / 3.0) * tr * tau_iso;
        if (a == b) jc += vol - Number(2.0 / 3.0) * tau_iso_src;
        Number geo = 0;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          const Number tau_mb = tau_bar[m][b] + (m == b ? dPsi_J - tr_tau_bar_dim : Number(0));
          geo += g[a][m] * tau_mb;
        }
        Pm[a][b] = (jc + geo) * inv_det;
      }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int f = 0; f < 3; ++f)
        grad[c][f] = d.jxw * (Pm[c][0] * d.ji[f * 3 + 0] + Pm[c][1] * d.ji[f * 3 + 1] + Pm[c][2] * d.ji[f * 3 + 2]);
      val[c] = d.jxw * rho_alpha * inv_det * val[c];
    }
  }
};

// Heat: value v, gradient c*dt*grad v (heat_transfer.h:242-248).
// Fields: Jinv (9), JxW, c per point; or, when every cell is affine (the
// affine-cell geometry of MatrixFree), Jinv and det J once per cell
// (cell_geo, 10 words), JxW = det J * wq[point], and c the only field.
template <typename Number>
struct QHeat {
  static constexpr int NC = 1;
  static constexpr int kFields = 11;  // 1 with cell_geo
  static constexpr bool kNodalField = false;
  const Number *qd;
  int64_t nqp;
  Number dt;
  const Number *cell_geo;  // per stored cell: Jinv (9), det J; or null
  const Number *wq;        // N^3 tensor-product quadrature weights (with cell_geo)
  struct Data {
    Number ji[9], jxw, c;
  };
  __device__ Data load(int64_t q, int cell, int pt) const {
    Data d;
    if (cell_geo) {
      const ConstantPtr<Number> g = as_constant(cell_geo) + (int64_t)cell * 10;
#pragma unroll
      for (int f = 0; f < 9; ++f) d.ji[f] = g[f];
      d.jxw = g[9] * wq[pt];
      d.c = stream_load(qd + q);
      return d;
    }
#pragma unroll
    for (int f = 0; f < 9; ++f) d.ji[f] = stream_load(qd + f * nqp + q);
    d.jxw = stream_load(qd + 9 * nqp + q);
    d.c = stream_load(qd + 10 * nqp + q);
    return d;
  }
  __device__ void apply(const Data &d, Number (&val)[1], Number (&grad)[1][3]) const {
    Number gx[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) gx[e] = grad[0][0] * d.ji[e] + grad[0][1] * d.ji[3 + e] + grad[0][2] * d.ji[6 + e];
    const Number s = d.c * dt * d.jxw;
#pragma unroll
    for (int f = 0; f < 3; ++f) grad[0][f] = s * (gx[0] * d.ji[f * 3] + gx[1] * d.ji[f * 3 + 1] + gx[2] * d.ji[f * 3 + 2]);
    val[0] = d.jxw * val[0];
  }
};

// The other formulation-1 cachings (MF caching parameter,
// mf_nh_operator.h:848-1010).  They define the same operator as tensor2 and
// trade memory for arithmetic differently.

// scalar_referential (mf_nh_operator.h:848-943): reference-frame geometry
// and one scalar per point; F is rebuilt from the cached nodal displacement
// at every application.  Fields: Jinv_ref (9), JxW_ref, mu - 2 lambda ln J.
template <typename Number>
struct QSolidSR {
  static constexpr int NC = 3;
  static constexpr int kFields = 11;
  static constexpr bool kNodalField = true;  // reference gradient of u at the point, set by attach()
  const Number *qd;
  int64_t nqp;
  const Number *u;  // nodal linearisation point, 3 per node
  Number mu, two_lambda, rho_alpha;
  struct Data {
    Number ji[9], jxw, sc, ug[9];
  };
  __device__ Data load(int64_t q, int cell, int pt) const {
    Data d;
#pragma unroll
    for (int f = 0; f < 9; ++f) d.ji[f] = stream_load(qd + f * nqp + q);
    d.jxw = stream_load(qd + 9 * nqp + q);
    d.sc = stream_load(qd + 10 * nqp + q);
    return d;
  }
  __device__ void attach(Data &d, const Number (&ug)[3][3]) const {
#pragma unroll
    for (int f = 0; f < 9; ++f) d.ug[f] = ug[f / 3][f % 3];
  }
  __device__ void apply(const Data &d, Number (&val)[3], Number (&grad)[3][3]) const {
    Number g[3][3], F[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        g[c][e] = grad[c][0] * d.ji[0 * 3 + e] + grad[c][1] * d.ji[1 * 3 + e] + grad[c][2] * d.ji[2 * 3 + e];
        F[c][e] = d.ug[c * 3 + 0] * d.ji[0 * 3 + e] + d.ug[c * 3 + 1] * d.ji[1 * 3 + e] +
                  d.ug[c * 3 + 2] * d.ji[2 * 3 + e] + (c == e ? Number(1) : Number(0));
      }
    const Number det = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
                       F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
    const Number id = Number(1) / det;
    Number Fi[3][3];
    Fi[0][0] = (F[1][1] * F[2][2] - F[1][2] * F[2][1]) * id;
    Fi[0][1] = (F[0][2] * F[2][1] - F[0][1] * F[2][2]) * id;
    Fi[0][2] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) * id;
    Fi[1][0] = (F[1][2] * F[2][0] - F[1][0] * F[2][2]) * id;
    Fi[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) * id;
    Fi[1][2] = (F[0][2] * F[1][0] - F[0][0] * F[1][2]) * id;
    Fi[2][0] = (F[1][0] * F[2][1] - F[1][1] * F[2][0]) * id;
    Fi[2][1] = (F[0][1] * F[2][0] - F[0][0] * F[2][1]) * id;
    Fi[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) * id;
    // spatial gradient of the test direction: grad_X v F^{-1}
    Number gx[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < 3; ++e) gx[c][e] = Fi[0][e] * g[c][0] + Fi[1][e] * g[c][1] + Fi[2][e] * g[c][2];
    const Number trg = gx[0][0] + gx[1][1] + gx[2][2];
    // tau = mu F F^T - (mu - 2 lambda ln J) I
    Number tau[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        tau[a][b] = mu * (F[a][0] * F[b][0] + F[a][1] * F[b][1] + F[a][2] * F[b][2]) - (a == b ? d.sc : Number(0));
    Number Q[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        Q[a][b] = d.sc * (gx[a][b] + gx[b][a]) + (a == b ? two_lambda * trg : Number(0)) + gx[a][0] * tau[0][b] +
                  gx[a][1] * tau[1][b] + gx[a][2] * tau[2][b];
    // first Piola form: (jc + gx tau) F^{-T}, then to reference gradients
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Number Pm[3];
#pragma unroll
      for (int b = 0; b < 3; ++b) Pm[b] = Q[c][0] * Fi[b][0] + Q[c][1] * Fi[b][1] + Q[c][2] * Fi[b][2];
#pragma unroll
      for (int f = 0; f < 3; ++f) grad[c][f] = d.jxw * (Pm[0] * d.ji[f * 3 + 0] + Pm[1] * d.ji[f * 3 + 1] + Pm[2] * d.ji[f * 3 + 2]);
      val[c] = d.jxw * rho_alpha * val[c];
    }
  }
};

// tensor4 (mf_nh_operator.h:976-994): deformed-frame geometry, 1/J, tau/J
// and the symmetric fourth-order tangent C = (2 s_c/J) S + (2 lambda/J) I x I
// as a 6 x 6 table over the symmetric index pairs (xx,yy,zz,xy,xz,yz).
// Fields: Jinv_cur (9), JxW_cur, 1/J, tau/J (6), C (36); the C entries are
// read while contracting.
template <typename Number>
struct QSolidT4 {
  static constexpr int NC = 3;
  static constexpr int kFields = 53;
  static constexpr bool kNodalField = false;
  const Number *qd;
  int64_t nqp;
  Number rho_alpha;
  struct Data {
    Number ji[9], jxw, c1, T[6];
    int64_t q;
  };
  __device__ Data load(int64_t q, int cell, int pt) const {
    Data d;
#pragma unroll
    for (int f = 0; f < 9; ++f) d.ji[f] = stream_load(qd + f * nqp + q);
    d.jxw = stream_load(qd + 9 * nqp + q);
    d.c1 = stream_load(qd + 10 * nqp + q);
#pragma unroll
    for (int f = 0; f < 6; ++f) d.T[f] = stream_load(qd + (11 + f) * nqp + q);
    d.q = q;
    return d;
  }
  __device__ void apply(const Data &d, Number (&val)[3], Number (&grad)[3][3]) const {
    Number g[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < 3; ++e) g[c][e] = grad[c][0] * d.ji[0 * 3 + e] + grad[c][1] * d.ji[1 * 3 + e] + grad[c][2] * d.ji[2 * 3 + e];
    // C : sym(g), summing both orders of the off-diagonal pairs
    Number sym2[6];
#pragma unroll
    for (int t = 0; t < 6; ++t)
      sym2[t] = (t < 3 ? Number(0.5) : Number(1)) * (g[kSymI[t]][kSymJ[t]] + g[kSymJ[t]][kSymI[t]]);
    Number Cs[6];
#pragma unroll
    for (int s6 = 0; s6 < 6; ++s6) {
      Number acc = 0;
#pragma unroll
      for (int t = 0; t < 6; ++t) acc += stream_load(qd + (17 + 6 * s6 + t) * nqp + d.q) * sym2[t];
      Cs[s6] = acc;
    }
    const Number Cm[3][3] = {{Cs[0], Cs[3], Cs[4]}, {Cs[3], Cs[1], Cs[5]}, {Cs[4], Cs[5], Cs[2]}};
    const Number Tm[3][3] = {{d.T[0], d.T[3], d.T[4]}, {d.T[3], d.T[1], d.T[5]}, {d.T[4], d.T[5], d.T[2]}};
    Number Pm[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) Pm[a][b] = Cm[a][b] + g[a][0] * Tm[0][b] + g[a][1] * Tm[1][b] + g[a][2] * Tm[2][b];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int f = 0; f < 3; ++f)
        grad[c][f] = d.jxw * (Pm[c][0] * d.ji[f * 3 + 0] + Pm[c][1] * d.ji[f * 3 + 1] + Pm[c][2] * d.ji[f * 3 + 2]);
      val[c] = d.jxw * rho_alpha * d.c1 * val[c];
    }
  }
};

// tensor4_ns (mf_nh_operator.h:996-1010): reference-frame geometry and the
// full (non-symmetric) fourth-order tangent A (81 entries), contracted as
// double_contract<2,0,3,1>(A, Grad v).  Fields: Jinv_ref (9), JxW_ref, A (81).
template <typename Number>
struct QSolidT4NS {
  static constexpr int NC = 3;
  static constexpr int kFields = 91;
  static constexpr bool kNodalField = false;
  const Number *qd;
  int64_t nqp;
  Number rho_alpha;
  struct Data {
    Number ji[9], jxw;
    int64_t q;
  };
  __device__ Data load(int64_t q, int cell, int pt) const {
    Data d;
#pragma unroll
    for (int f = 0; f < 9; ++f) d.ji[f] = stream_load(qd + f * nqp + q);
    d.jxw = stream_load(qd + 9 * nqp + q);
    d.q = q;
    return d;
  }
  __device__ void apply(const Data &d, Number (&val)[3], Number (&grad)[3][3]) const {
    Number g[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < 3; ++e) g[c][e] = grad[c][0] * d.ji[0 * 3 + e] + grad[c][1] * d.ji[1 * 3 + e] + grad[c][2] * d.ji[2 * 3 + e];
    Number Pm[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        Number acc = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int l = 0; l < 3; ++l) acc += stream_load(qd + (10 + ((a * 3 + b) * 3 + k) * 3 + l) * nqp + d.q) * g[k][l];
        Pm[a][b] = acc;
      }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int f = 0; f < 3; ++f)
        grad[c][f] = d.jxw * (Pm[c][0] * d.ji[f * 3 + 0] + Pm[c][1] * d.ji[f * 3 + 1] + Pm[c][2] * d.ji[f * 3 + 2]);
      val[c] = d.jxw * rho_alpha * val[c];
    }
  }
};

// ------------------------------------------------------ apply kernel ---
// One launch covers the cells [cell_begin, cell_end) of one colour: no two of
// them share a node, so the scatter is a plain read-modify-write and the
// result is bit-reproducible (cells are stored colour-major).
// fix_mode: 0 no constrained-row handling, 1 dst[c] = src[c], 2 dst[c] += src[c]
template <typename Number>
struct ApplyArgs {
  const uint32_t *cell_nodes;
  int cell_begin, cell_end;
  const Number *src;
  Number *dst;
  const Number *shape;   // N x N
  const Number *colloc;  // N x N
  const int32_t *constrained;
  int64_t n_constrained;
  int fix_mode;
};

// The cells [cell_begin, min(cell_begin + cpb, cell_end)) on all threads of
// the workgroup, thread t owning point t % N3 of local cell t / N3; threads
// past cpb*N3 only join the barriers.  `assign`: the first colour touching a
// node stores instead of adding.  Shape and collocation tables must be in LDS
// before the first barrier inside.  kToCells: instead of the read-modify-write
// of dst, each point stores its node's contribution (NC values) at
// dst[(cell N3 + point) NC + c], for the gather of the GATHER schedule.
// The NC values of one node, loaded as one struct so that the compiler can use
// wide loads (nodes are 8-byte aligned; the hardware takes unaligned wide loads).
template <typename Number, int NC>
struct NodeValues {
  Number v[NC];
};
template <typename Number, int NC>
__device__ inline NodeValues<Number, NC> load_node(const Number *p) {
  return *reinterpret_cast<const NodeValues<Number, NC> *>(p);
}
template <typename Number, int NC>
__device__ inline void store_node(Number *p, const Number (&v)[NC]) {
  NodeValues<Number, NC> x;
#pragma unroll
  for (int c = 0; c < NC; ++c) x.v[c] = v[c];
  *reinterpret_cast<NodeValues<Number, NC> *>(p) = x;
}

// g_shape / g_colloc: the N x N 1D matrices in global memory, copied into
// s_shape / s_colloc here.  The LDS tile and the two tables must not be in use
// by other threads of the workgroup on entry (a caller that runs several
// groups in sequence puts a barrier between them).
template <int P, typename Number, class QOp, bool kToCells = false, bool kSpareSlots = false>
__device__ inline void cell_group(const QOp &qop, const uint32_t *__restrict__ cell_nodes,
                                  const Number *__restrict__ src, Number *__restrict__ dst, int cell_begin,
                                  int cell_end, int cpb, int assign, const Number *__restrict__ g_shape,
                                  const Number *__restrict__ g_colloc, Number *s_shape, Number *s_colloc,
                                  Number *s_tile) {
  constexpr int N = P + 1, N3 = N * N * N, NC = QOp::NC;
  const int t = threadIdx.x;
  const int lc_raw = t / N3;
  const bool pad = lc_raw >= cpb;
  const int lc = pad ? cpb - 1 : lc_raw;
  const int pt = pad ? 0 : t - lc_raw * N3;
  Tile<N, Number> T;
  T.base = s_tile + lc * (kSpareSlots ? 5 : 3) * NC * N3;
  T.pt = pt;
  T.i = pt % N;
  T.j = (pt / N) % N;
  T.k = pt / (N * N);
  T.writer = !pad;
  const int cell = cell_begin + lc;
  const bool active = !pad && cell < cell_end;
  const int64_t qidx = active ? (int64_t)cell * N3 + pt : 0;

  uint32_t entry = active ? cell_nodes[qidx] : 0u;
  // the 1D matrices are loaded with the map and written to LDS just before
  // the first barrier of the sweeps, which publishes them: no wait on these
  // loads stands before the gather
  const bool table_lane = t < N * N;
  Number sh = 0, co = 0;
  if (table_lane) {
    sh = g_shape[t];
    co = g_colloc[t];
  }
  const int64_t node = entry & kNodeMaskDev;
  // a node's NC values are contiguous: one wide load (FP64, NC = 3: a 16- and
  // an 8-byte load instead of three) for all components, constrained ones
  // dropped afterwards
  Number u[NC];
  {
    const NodeValues<Number, NC> v = active ? load_node<Number, NC>(src + node * NC) : NodeValues<Number, NC>{};
#pragma unroll
    for (int c = 0; c < NC; ++c) u[c] = ((entry >> (kConstraintShift + c)) & 1u) ? Number(0) : v.v[c];
  }
  // dst of this node: no other workgroup of the launch writes it, so the
  // scatter at the end is a plain store.  The first colour touching a node
  // stores without reading.  Read with the gather, the load's latency hides
  // behind the sweeps; read after them, its 2 NC registers are free during
  // the sweeps.  For the FP64 solid ops with per-point tensors and p <= 4 the
  // late read gives the kernel one more wave per SIMD and is the faster of the
  // two (measured); for the heat op and the 10-word tensor2 op, whose wave
  // count it does not change, and elsewhere the early read is.
  constexpr bool kLateDst = sizeof(Number) == 8 && P <= 4 && !std::is_same<QOp, QHeat<Number>>::value &&
                            !std::is_same<QOp, QSolidT2F1G<Number>>::value &&
                            !std::is_same<QOp, QSolidT2F1G<Number, true>>::value &&
                            !std::is_same<QOp, QSolidT2F1G<Number, true, true>>::value;
  const bool first = assign && (entry >> 31);
  Number dold[NC];
  if constexpr (!kLateDst && !kToCells) {
    // constrained components are read but not used: their dst entries are not stored
    const NodeValues<Number, NC> v =
        active && !first ? load_node<Number, NC>(dst + node * NC) : NodeValues<Number, NC>{};
#pragma unroll
    for (int c = 0; c < NC; ++c) dold[c] = v.v[c];
  }
  // the cached nodal field of the nodal-field ops, gathered with src
  NodeValues<Number, 3> w{};
  if constexpr (QOp::kNodalField) {
    if (active) w = load_node<Number, 3>(qop.u + node * 3);
  }
  // one cell per workgroup: the cell index is the same on every lane (padding
  // lanes included), which lets the per-cell tables come in as scalar loads
  const int qcell = cpb == 1 ? __builtin_amdgcn_readfirstlane(cell < cell_end ? cell : 0) : (active ? cell : 0);
  typename QOp::Data qd = qop.load(qidx, qcell, pt);
  // read only after the barrier that follows the first puts of the sweeps
  if (table_lane) {
    s_shape[t] = sh;
    s_colloc[t] = co;
  }
  if constexpr (QOp::kNodalField) {
    // reference gradient of the cached nodal field (read plain: constrained
    // entries as stored), mf_nh_operator.h:873-877
    Number wg[3][3];
    evaluate_g<N, 3, Number>(T, s_shape, s_colloc, w.v, wg);
    qop.attach(qd, wg);
  }

  Number val[NC], grad[NC][3];
  evaluate_vg<N, NC, Number>(T, s_shape, s_colloc, u, val, grad);
  qop.apply(qd, val, grad);
  Number out[NC];
  integrate_vg<N, NC, Number, kSpareSlots>(T, s_shape, s_colloc, val, grad, out);
  if constexpr (kToCells) {
    if (active) store_node<Number, NC>(dst + qidx * NC, out);
    return;
  }
  if constexpr (kLateDst) {
    const NodeValues<Number, NC> v =
        active && !first ? load_node<Number, NC>(dst + node * NC) : NodeValues<Number, NC>{};
#pragma unroll
    for (int c = 0; c < NC; ++c) dold[c] = v.v[c];
  }
  if (active) {
    Number r[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) r[c] = dold[c] + out[c];
    if (((entry >> kConstraintShift) & ((1u << NC) - 1u)) == 0u) {
      store_node<Number, NC>(dst + node * NC, r);  // no constrained component: one wide store
    } else {
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (!((entry >> (kConstraintShift + c)) & 1u)) dst[node * NC + c] = r[c];
    }
  }
}

// Spare tile slots (5 NC instead of 3 NC per cell, two barriers fewer per
// cell) where the larger tile stays within 16 KB per workgroup.
template <int P, int CPB, typename Number, int NC>
constexpr bool spare_slots() {
  return CPB * 5 * NC * (P + 1) * (P + 1) * (P + 1) * (int)sizeof(Number) <= 16384;
}

template <int P, int CPB, typename Number, class QOp>
__global__ void __launch_bounds__((block_threads<P, CPB>()))
    apply_cells(ApplyArgs<Number> a, QOp qop, int assign) {
  constexpr int N = P + 1, N3 = N * N * N, NC = QOp::NC, NT = block_threads<P, CPB>();
  constexpr bool kSpare = spare_slots<P, CPB, Number, NC>();
  __shared__ Number s_shape[N * N], s_colloc[N * N];
  __shared__ Number s_tile[CPB * (kSpare ? 5 : 3) * NC * N3];
  const int t = threadIdx.x;
  const int first_cell = a.cell_begin + xcd_remap(blockIdx.x, gridDim.x) * CPB;
  cell_group<P, Number, QOp, false, kSpare>(qop, a.cell_nodes, a.src, a.dst, first_cell, a.cell_end, CPB, assign,
                                            a.shape, a.colloc, s_shape, s_colloc, s_tile);
  // rows of constrained dofs: identity (mf_nh_operator.h:622-624).  The
  // cells never store a constrained component, so these rows can follow them.
  if (a.fix_mode) {
    const int64_t g = (int64_t)blockIdx.x * NT + t;
    for (int64_t idx = g; idx < a.n_constrained; idx += (int64_t)gridDim.x * NT) {
      const int64_t dof = a.constrained[idx];
      if (a.fix_mode == 1) a.dst[dof] = a.src[dof];
      else a.dst[dof] += a.src[dof];
    }
  }
}

// GATHER schedule, first launch: every cell at once, each point storing its
// contribution per cell (no colours, no read of dst).
template <int P, int CPB, typename Number, class QOp>
__global__ void __launch_bounds__((block_threads<P, CPB>()))
    apply_cells_to_cells(ApplyArgs<Number> a, QOp qop, Number *contrib) {
  constexpr int N = P + 1, N3 = N * N * N, NC = QOp::NC;
  constexpr bool kSpare = spare_slots<P, CPB, Number, NC>();  // as apply_cells: the same arithmetic
  __shared__ Number s_shape[N * N], s_colloc[N * N];
  __shared__ Number s_tile[CPB * (kSpare ? 5 : 3) * NC * N3];
  const int first_cell = a.cell_begin + xcd_remap(blockIdx.x, gridDim.x) * CPB;
  cell_group<P, Number, QOp, true, kSpare>(qop, a.cell_nodes, a.src, contrib, first_cell, a.cell_end, CPB, 1, a.shape,
                                           a.colloc, s_shape, s_colloc, s_tile);
}

// GATHER schedule, second launch: one thread per node sums the contributions
// of its cells in colour order -- the order in which the colour launches add
// them, so the result is theirs bit for bit (up to the sign of a zero).
// fix_mode 1 (vmult): constrained components copy src, the others are
// overwritten; 2 (vmult_add): constrained components add src, the others add
// to dst.
template <typename Number, int NC>
__device__ inline void gather_node(const Number *__restrict__ contrib, const int32_t *__restrict__ offsets,
                                   const int32_t *__restrict__ points, const uint8_t *__restrict__ cmask,
                                   const Number *__restrict__ src, Number *__restrict__ dst, int64_t node,
                                   int fix_mode) {
  const int b = offsets[node], e = offsets[node + 1];
  const unsigned m = cmask[node];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int64_t dof = node * NC + c;
    if ((m >> c) & 1u) {
      dst[dof] = fix_mode == 1 ? src[dof] : dst[dof] + src[dof];
      continue;
    }
    int k = b;
    Number acc;
    if (fix_mode == 1) {
      acc = contrib[(int64_t)points[k] * NC + c];
      ++k;
    } else {
      acc = dst[dof];
    }
    for (; k < e; ++k) acc = acc + contrib[(int64_t)points[k] * NC + c];
    dst[dof] = acc;
  }
}

template <typename Number, int NC>
__global__ void gather_contributions(const Number *__restrict__ contrib, const int32_t *__restrict__ offsets,
                                     const int32_t *__restrict__ points, const uint8_t *__restrict__ cmask,
                                     const Number *__restrict__ src, Number *__restrict__ dst, int64_t n_nodes,
                                     int fix_mode) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= n_nodes) return;
  gather_node<Number, NC>(contrib, offsets, points, cmask, src, dst, node, fix_mode);
}

}  // namespace mfg
