// Device operators behind the C ABI: the neo-Hookean tangent operator
// (NeoHookOperator, mf_nh_operator.h) and the heat operator (LaplaceOperator,
// heat_transfer.h:81-250).  Owns the device copy of the level layout, the
// quadrature cache and the diagonal; vectors are caller-owned device memory.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "operator.hpp"
#include "setup_kernels.hpp"
#include "vector_kernels.hpp"

using namespace mfg;

namespace mfg {

// per-cell reference JxW of the tensor2 formulation-1 op, or null when it is
// a per-point field
template <typename Number>
const Number *t2_cell_jxw(const OperatorData &op) {
  return op.affine_ref ? op.cell_jxw.as<Number>() : nullptr;
}

// tensor2 formulation 1 from the 10-word storage (affine meshes only), with
// the constants of material m
template <typename Number, bool kEarlyCell = false, bool kCartesian = false>
QSolidT2F1G<Number, kEarlyCell, kCartesian> t2_gradient_op(const OperatorData &op, int m = 0) {
  const OperatorData::MaterialConstants &M = op.mats[m];
  return QSolidT2F1G<Number, kEarlyCell, kCartesian>{op.qd.as<Number>(),       op.nqp,
                                                     Number(M.mu),             Number(2.0 * M.lambda),
                                                     Number(M.rho_alpha),      op.cell_jac.as<Number>(),
                                                     op.cell_jxw.as<Number>(), op.wq3.as<Number>()};
}

// per-cell geometry of the heat op on affine meshes, or null (per-point fields)
template <typename Number>
const Number *heat_cell_geo(const OperatorData &op) {
  return op.affine_ref ? op.cell_geo.as<Number>() : nullptr;
}

// fn(q) with the quadrature op of this operator for the cells of material m
// (the heat op has one).  kEarlyCell: one cell per workgroup, the 10-word op
// takes its per-cell tables into SGPRs (only their diagonals when every
// cell's reference Jacobian is diagonal).
template <typename Number, bool kEarlyCell, class F>
static void with_qop(const OperatorData &op, int m, F fn) {
  if (op.kind == 1) {
    fn(QHeat<Number>{op.qd.as<Number>(), op.nqp, Number(op.dt), heat_cell_geo<Number>(op), op.wq3.as<Number>()});
    return;
  }
  const OperatorData::MaterialConstants &M = op.mats[m];
  const Number ra = Number(M.rho_alpha);
  if (op.sp.caching == MFG_CACHING_SCALAR_REFERENTIAL) {
    fn(QSolidSR<Number>{op.qd.as<Number>(), op.nqp, op.u_nodal.as<Number>(), Number(M.mu), Number(2.0 * M.lambda), ra});
  } else if (op.sp.caching == MFG_CACHING_TENSOR4) {
    fn(QSolidT4<Number>{op.qd.as<Number>(), op.nqp, ra});
  } else if (op.sp.caching == MFG_CACHING_TENSOR4_NS) {
    fn(QSolidT4NS<Number>{op.qd.as<Number>(), op.nqp, ra});
  } else if (op.t2_gradient()) {
    if constexpr (kEarlyCell) {
      if (op.cartesian_ref) {
        fn(t2_gradient_op<Number, true, true>(op, m));
        return;
      }
    }
    fn(t2_gradient_op<Number, kEarlyCell>(op, m));
  } else if (op.sp.formulation == 1) {
    fn(QSolidT2F1<Number>{op.qd.as<Number>(), op.nqp, Number(2.0 * M.lambda), ra, t2_cell_jxw<Number>(op),
                          op.wq3.as<Number>()});
  } else {
    fn(QSolidT2F0<Number>{op.qd.as<Number>(), op.nqp, Number(M.kappa), Number(M.c_1), ra});
  }
}

template <typename Number>
static void fill(Number *x, int64_t n, Number v, hipStream_t s) {
  if (n <= 0) return;
  const int threads = 256;
  const int blocks = (int)std::min<int64_t>((n + threads - 1) / threads, 65536);
  fill_kernel<Number><<<blocks, threads, 0, s>>>(x, n, v);
  MFG_HIP(hipGetLastError());
}

// --------------------------------------------------------- dispatchers ---
template <int P, typename Number>
static void launch_apply(const OperatorData &op, const void *src, void *dst, int fix_mode, hipStream_t stream) {
  constexpr int CPB = DefaultCPB<P, Number>::value;
  ApplyArgs<Number> a;
  a.cell_nodes = op.cell_nodes.as<uint32_t>();
  a.src = (const Number *)src;
  a.dst = (Number *)dst;
  a.shape = op.tab_shape.as<Number>();
  a.colloc = op.tab_colloc.as<Number>();
  a.constrained = op.constrained.as<int32_t>();
  a.n_constrained = op.n_constrained;
  constexpr int NT = block_threads<P, CPB>();
  const int assign = fix_mode == 1 ? 1 : 0;  // vmult: first touch stores, no zero fill
  if (op.cell_schedule != MFG_CELLS_GATHER) {
    // one workgroup per group of CPB cells, one launch per colour (and
    // material); the first launch also writes the identity rows
    bool fix_pending = true;
    op.for_each_range([&](int, int b, int e, int m) {
      with_qop<Number, CPB == 1>(op, m, [&](auto q) {
        a.cell_begin = b;
        a.cell_end = e;
        a.fix_mode = fix_pending ? fix_mode : 0;
        fix_pending = false;
        apply_cells<P, CPB, Number, decltype(q)><<<blocks_for(e - b, CPB), NT, 0, stream>>>(a, q, assign);
        MFG_HIP(hipGetLastError());
      });
    });
    return;
  }
  // GATHER: all cells (of one material per launch) into the contribution
  // buffer, then the ordered per-node sums
  a.fix_mode = 0;
  auto cells = [&](int b, int e, int m) {
    with_qop<Number, CPB == 1>(op, m, [&](auto q) {
      a.cell_begin = b;
      a.cell_end = e;
      apply_cells_to_cells<P, CPB, Number, decltype(q)><<<blocks_for(e - b, CPB), NT, 0, stream>>>(
          a, q, op.contrib.as<Number>());
      MFG_HIP(hipGetLastError());
    });
  };
  if (op.has_inclusion()) op.for_each_range([&](int, int b, int e, int m) { cells(b, e, m); });
  else cells(0, op.n_cells, 0);
  const int64_t nn = op.n_nodes;
  auto sums = [&](auto nc_tag) {
    constexpr int NCv = decltype(nc_tag)::value;
    gather_contributions<Number, NCv><<<(int)((nn + 255) / 256), 256, 0, stream>>>(
        op.contrib.as<Number>(), op.gather_offsets.as<int32_t>(), op.gather_points.as<int32_t>(),
        op.gather_cmask.as<uint8_t>(), (const Number *)src, (Number *)dst, nn, fix_mode);
    MFG_HIP(hipGetLastError());
  };
  if (op.nc == 3) sums(std::integral_constant<int, 3>());
  else sums(std::integral_constant<int, 1>());
}

template <int P, typename Number>
static void launch_cache(OperatorData &op, const void *u, int u_fp32) {
  constexpr int CPB = DefaultCPB<P, Number>::value;
  CacheArgs<Number> a;
  a.cell_nodes = op.cell_nodes.as<uint32_t>();
  a.u = u;
  a.u_fp32 = u_fp32;
  a.vertices = op.vertices.as<double>();
  a.shape = op.tab_shape.as<Number>();
  a.colloc = op.tab_colloc.as<Number>();
  a.points = op.tab_points.as<double>();
  a.weights = op.tab_weights.as<double>();
  a.qd = op.qd.as<Number>();
  a.nqp = op.nqp;
  a.formulation = op.sp.formulation;
  a.caching = op.sp.caching;
  a.jxw_ref_field = op.affine_ref ? 0 : 1;
  a.t2_gradient = op.t2_gradient() ? 1 : 0;
  a.error = op.flag.as<int>();
  auto cells = [&](int b, int e, int m) {
    a.cell_begin = b;
    a.cell_end = e;
    a.mu = Number(op.mats[m].mu);
    a.lambda = Number(op.mats[m].lambda);
    cache_solid<P, CPB, Number><<<blocks_for(e - b, CPB), block_threads<P, CPB>(), 0, op.stream>>>(a);
    MFG_HIP(hipGetLastError());
  };
  if (op.has_inclusion()) op.for_each_range([&](int, int b, int e, int m) { cells(b, e, m); });
  else cells(0, op.n_cells, 0);
}

template <int P, typename Number>
static void launch_diagonal(OperatorData &op) {
  DiagArgs<Number> a;
  a.cell_nodes = op.cell_nodes.as<uint32_t>();
  a.sq = op.tab_sq.as<Number>();
  a.shape = op.tab_shape.as<Number>();
  a.colloc = op.tab_colloc.as<Number>();
  a.diag = op.diag.as<Number>();
  constexpr int NT = ((P + 1) * (P + 1) * (P + 1) + 63) / 64 * 64;
  op.for_each_range([&](int, int b, int e, int m) {
    with_qop<Number, false>(op, m, [&](auto q) {
      a.cell_begin = b;
      a.cell_end = e;
      diagonal_cells<P, Number, decltype(q)><<<e - b, NT, 0, op.stream>>>(a, q);
      MFG_HIP(hipGetLastError());
    });
  });
}

#define MFG_DISPATCH_DEGREE(p, FN, ...)                                                  \
  switch (p) {                                                                           \
    case 1: FN<1, Number>(__VA_ARGS__); break;                                           \
    case 2: FN<2, Number>(__VA_ARGS__); break;                                           \
    case 3: FN<3, Number>(__VA_ARGS__); break;                                           \
    case 4: FN<4, Number>(__VA_ARGS__); break;                                           \
    case 5: FN<5, Number>(__VA_ARGS__); break;                                           \
    case 6: FN<6, Number>(__VA_ARGS__); break;                                           \
    default: throw Error(MFG_ERR_ARGUMENT, "unsupported degree");                        \
  }

template <typename Number>
static void apply_t(const OperatorData &op, const void *src, void *dst, int fix_mode, hipStream_t stream) {
  MFG_DISPATCH_DEGREE(op.p, launch_apply, op, src, dst, fix_mode, stream);
}
template <typename Number>
static void cache_t(OperatorData &op, const void *u, int u_fp32) {
  MFG_DISPATCH_DEGREE(op.p, launch_cache, op, u, u_fp32);
}
template <typename Number>
static void diagonal_t(OperatorData &op) {
  MFG_DISPATCH_DEGREE(op.p, launch_diagonal, op);
}

void OperatorData::apply(const void *src, void *dst, bool add, hipStream_t on) const {
  MFG_REQUIRE(ready, MFG_ERR_STATE, "operator used before cache()");
  MFG_REQUIRE(src && dst, MFG_ERR_ARGUMENT, "null vector");
  // the scatter reads dst while other workgroups still gather src
  MFG_REQUIRE(src != dst, MFG_ERR_ARGUMENT, "dst and src must be different vectors");
  const hipStream_t s = on ? on : stream;
  // vmult needs no zero fill: the first colour touching a node stores
  timing_begin(s);
  if (precision == MFG_FP64) apply_t<double>(*this, src, dst, add ? 2 : 1, s);
  else apply_t<float>(*this, src, dst, add ? 2 : 1, s);
  timing_end(s);
}

// Event pairs are reused; at most kMaxEvents launches are kept between
// queries (older ones are folded in by an early collect).
static constexpr size_t kMaxEvents = 2048;

void OperatorData::timing_begin(hipStream_t s) const {
  if (!timing) return;
  if (events_used == events.size()) {
    if (events.size() >= kMaxEvents) throw Error(MFG_ERR_STATE, "too many timed launches between queries");
    hipEvent_t a, b;
    MFG_HIP(hipEventCreate(&a));
    MFG_HIP(hipEventCreate(&b));
    events.emplace_back(a, b);
  }
  MFG_HIP(hipEventRecord(events[events_used].first, s));
}

void OperatorData::timing_end(hipStream_t s) const {
  if (!timing) return;
  MFG_HIP(hipEventRecord(events[events_used].second, s));
  ++events_used;
}

void OperatorData::timing_collect(double &ms, int64_t &n) {
  ms = 0;
  n = (int64_t)events_used;
  for (size_t i = 0; i < events_used; ++i) {
    MFG_HIP(hipEventSynchronize(events[i].second));
    float t = 0;
    MFG_HIP(hipEventElapsedTime(&t, events[i].first, events[i].second));
    ms += t;
  }
  events_used = 0;
}

OperatorData::~OperatorData() {
  for (auto &e : events) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
}

void OperatorData::compute_diagonal() {
  MFG_REQUIRE(ready, MFG_ERR_STATE, "compute_diagonal before cache()");
  const size_t es = precision == MFG_FP64 ? 8 : 4;
  if (diag.bytes != n_dofs * es || inv_diag.bytes != n_dofs * es) {
    diag.allocate(n_dofs * es);
    inv_diag.allocate(n_dofs * es);
    ++buffers_version;
  }
  fill<int>(flag.as<int>(), 1, 0, stream);
  const int threads = 256;
  if (precision == MFG_FP64) {
    fill<double>(diag.as<double>(), n_dofs, 0.0, stream);
    diagonal_t<double>(*this);
    if (n_constrained)
      set_entries_kernel<double><<<blocks_for(n_constrained, threads), threads, 0, stream>>>(
          diag.as<double>(), constrained.as<int32_t>(), n_constrained, 1.0);
    invert_kernel<double><<<std::min(blocks_for(n_dofs, threads), 65536), threads, 0, stream>>>(
        diag.as<double>(), inv_diag.as<double>(), n_dofs, flag.as<int>());
  } else {
    fill<float>(diag.as<float>(), n_dofs, 0.0f, stream);
    diagonal_t<float>(*this);
    if (n_constrained)
      set_entries_kernel<float><<<blocks_for(n_constrained, threads), threads, 0, stream>>>(
          diag.as<float>(), constrained.as<int32_t>(), n_constrained, 1.0f);
    invert_kernel<float><<<std::min(blocks_for(n_dofs, threads), 65536), threads, 0, stream>>>(
        diag.as<float>(), inv_diag.as<float>(), n_dofs, flag.as<int>());
  }
  MFG_HIP(hipGetLastError());
  int h = 0;
  MFG_HIP(hipMemcpyAsync(&h, flag.ptr, sizeof(int), hipMemcpyDeviceToHost, stream));
  MFG_HIP(hipStreamSynchronize(stream));
  MFG_REQUIRE(h == 0, MFG_ERR_NUMERIC, "No diagonal entry in a positive definite operator should be zero or negative");
  has_diagonal = true;
}

void OperatorData::cache(const void *u, int u_precision) {
  MFG_REQUIRE(kind == 0, MFG_ERR_ARGUMENT, "cache() applies to the solid operator");
  MFG_REQUIRE(u, MFG_ERR_ARGUMENT, "null displacement");
  fill<int>(flag.as<int>(), 1, 0, stream);
  if (sp.caching == MFG_CACHING_SCALAR_REFERENTIAL) {  // keeps the displacement itself (mf_nh_operator.h:442-455)
    const int threads = 256, blocks = std::min(blocks_for(n_dofs, threads), 65536);
    if (u_nodal.bytes != (size_t)n_dofs * elem_size()) {
      u_nodal.allocate((size_t)n_dofs * elem_size());
      ++buffers_version;
    }
    if (precision == MFG_FP64 && u_precision == MFG_FP64)
      convert_kernel<double, double><<<blocks, threads, 0, stream>>>(u_nodal.as<double>(), (const double *)u, n_dofs);
    else if (precision == MFG_FP64)
      convert_kernel<double, float><<<blocks, threads, 0, stream>>>(u_nodal.as<double>(), (const float *)u, n_dofs);
    else if (u_precision == MFG_FP64)
      convert_kernel<float, double><<<blocks, threads, 0, stream>>>(u_nodal.as<float>(), (const double *)u, n_dofs);
    else
      convert_kernel<float, float><<<blocks, threads, 0, stream>>>(u_nodal.as<float>(), (const float *)u, n_dofs);
    MFG_HIP(hipGetLastError());
  }
  if (precision == MFG_FP64) cache_t<double>(*this, u, u_precision == MFG_FP32);
  else cache_t<float>(*this, u, u_precision == MFG_FP32);
  int h = 0;
  MFG_HIP(hipMemcpyAsync(&h, flag.ptr, sizeof(int), hipMemcpyDeviceToHost, stream));
  MFG_HIP(hipStreamSynchronize(stream));
  MFG_REQUIRE(h == 0, MFG_ERR_NUMERIC, "det_F is not positive");
  ready = true;
  has_diagonal = false;
}

template <typename Number>
static void upload_tables(OperatorData &op, const Tables1D &t) {
  const int N = t.N;
  std::vector<Number> sh(N * N), co(N * N), sq(3 * N * N);
  for (int x = 0; x < N * N; ++x) {
    sh[x] = (Number)t.shape[x];
    co[x] = (Number)t.colloc[x];
  }
  // derivative of the GLL basis at the Gauss points: colloc * shape
  for (int q = 0; q < N; ++q)
    for (int a = 0; a < N; ++a) {
      double s = t.shape[q * N + a], d = 0;
      for (int r = 0; r < N; ++r) d += t.colloc[q * N + r] * t.shape[r * N + a];
      sq[0 * N * N + q * N + a] = (Number)(s * s);
      sq[1 * N * N + q * N + a] = (Number)(s * d);
      sq[2 * N * N + q * N + a] = (Number)(d * d);
    }
  upload(op.tab_shape, sh.data(), sh.size(), op.stream);
  upload(op.tab_colloc, co.data(), co.size(), op.stream);
  upload(op.tab_sq, sq.data(), sq.size(), op.stream);
  upload(op.tab_points, t.points.data(), t.points.size(), op.stream);
  upload(op.tab_weights, t.weight.data(), t.weight.size(), op.stream);
}

// Greedy colouring of the cells in their given order such that no two cells
// of one colour share a node; on a lexicographic hex grid this yields the
// 8 parity colours.  Fills perm (colour-major order -> given cell index) and
// colour_start; with an inclusion material (solid operators with
// inclusion_mu != 0 and a cell whose material id is not 0) also mat_split,
// the material-0 cells of each colour coming first.
void OperatorData::colour_cells(const mfg_level_desc &d, int N3) {
  std::vector<int64_t> off((size_t)n_nodes + 1, 0);
  const int64_t ne = (int64_t)n_cells * N3;
  for (int64_t x = 0; x < ne; ++x) off[d.cell_nodes[x] + 1]++;
  for (int64_t i = 0; i < n_nodes; ++i) off[i + 1] += off[i];
  std::vector<int32_t> cells((size_t)ne);
  {
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    for (int64_t x = 0; x < ne; ++x) cells[pos[d.cell_nodes[x]]++] = (int32_t)(x / N3);
  }
  std::vector<int8_t> colour(n_cells, -1);
  int n_col = 0;
  for (int c = 0; c < n_cells; ++c) {
    uint64_t used = 0;
    for (int x = 0; x < N3; ++x) {
      const int32_t node = d.cell_nodes[(int64_t)c * N3 + x];
      for (int64_t k = off[node]; k < off[node + 1]; ++k) {
        const int8_t cc = colour[cells[k]];
        if (cc >= 0) used |= (uint64_t)1 << cc;
      }
    }
    const int cc = __builtin_ctzll(~used);
    MFG_REQUIRE(cc < 64, MFG_ERR_ARGUMENT, "cell colouring needs more than 64 colours");
    colour[c] = (int8_t)cc;
    n_col = std::max(n_col, cc + 1);
  }
  colour_start.assign(n_col + 1, 0);
  for (int c = 0; c < n_cells; ++c) colour_start[colour[c] + 1]++;
  for (int k = 0; k < n_col; ++k) colour_start[k + 1] += colour_start[k];
  perm.assign(n_cells, 0);
  mat_split.clear();
  const int32_t *ids = kind == 0 && sp.inclusion_mu != 0 ? d.material_ids : nullptr;
  const bool split = ids && std::any_of(ids, ids + n_cells, [](int32_t x) { return x != 0; });
  std::vector<int> fillp(colour_start.begin(), colour_start.end() - 1);
  if (!split) {
    for (int c = 0; c < n_cells; ++c) perm[fillp[colour[c]]++] = c;
    return;
  }
  mat_split.assign(colour_start.begin(), colour_start.end() - 1);
  for (int c = 0; c < n_cells; ++c)
    if (ids[c] == 0) ++mat_split[colour[c]];
  std::vector<int> fill1(mat_split);
  for (int c = 0; c < n_cells; ++c) perm[ids[c] == 0 ? fillp[colour[c]]++ : fill1[colour[c]]++] = c;
}

void OperatorData::build_gather() {
  MFG_REQUIRE(nqp < ((int64_t)1 << 31), MFG_ERR_ARGUMENT, "too many quadrature points for the GATHER tables");
  const int N3 = (p + 1) * (p + 1) * (p + 1);
  std::vector<int32_t> offsets((size_t)n_nodes + 1, 0), points((size_t)nqp);
  for (int64_t x = 0; x < nqp; ++x) offsets[host_cell_nodes[x] + 1]++;
  for (int64_t v = 0; v < n_nodes; ++v) offsets[v + 1] += offsets[v];
  std::vector<int32_t> fill(offsets.begin(), offsets.end() - 1);
  // stored (colour-major) cell order is the order of the colour launches
  for (int c = 0; c < n_cells; ++c)
    for (int pt = 0; pt < N3; ++pt) {
      const int32_t node = host_cell_nodes[(int64_t)perm[c] * N3 + pt];
      points[fill[node]++] = c * N3 + pt;
    }
  upload(gather_offsets, offsets.data(), offsets.size(), stream);
  upload(gather_points, points.data(), points.size(), stream);
  upload(gather_cmask, host_cmask.data(), host_cmask.size(), stream);
  contrib.allocate((size_t)nqp * nc * elem_size());
  ++buffers_version;
}

void OperatorData::setup_layout(const mfg_level_desc &d) {
  MFG_REQUIRE(d.degree >= 1 && d.degree <= kMaxDegree, MFG_ERR_ARGUMENT, "degree must be in 1..6");
  MFG_REQUIRE(d.n_q_1d == d.degree + 1, MFG_ERR_ARGUMENT, "n_q_1d must equal degree+1");
  MFG_REQUIRE(d.n_components == (kind == 0 ? 3 : 1), MFG_ERR_ARGUMENT, "wrong number of components");
  MFG_REQUIRE(d.n_cells > 0 && d.cell_nodes && d.cell_vertices, MFG_ERR_ARGUMENT, "empty level");
  MFG_REQUIRE(d.n_nodes > 0 && d.n_nodes <= (int64_t)kNodeMask, MFG_ERR_ARGUMENT, "node count out of range");
  MFG_REQUIRE(d.n_constrained == 0 || d.constrained_dofs, MFG_ERR_ARGUMENT, "null constraint list");
  p = d.degree;
  nc = d.n_components;
  n_cells = d.n_cells;
  n_nodes = d.n_nodes;
  n_dofs = n_nodes * nc;
  const int N3 = (p + 1) * (p + 1) * (p + 1);
  nqp = (int64_t)n_cells * N3;
  n_constrained = d.n_constrained;
  // constraint bits per node (component c -> bit 28+c)
  std::vector<uint8_t> cmask(n_nodes, 0);
  for (int64_t x = 0; x < d.n_constrained; ++x) {
    const int64_t dof = d.constrained_dofs[x];
    MFG_REQUIRE(dof >= 0 && dof < n_dofs, MFG_ERR_ARGUMENT, "constrained dof out of range");
    cmask[dof / nc] |= (uint8_t)(1u << (dof % nc));
  }
  for (int64_t x = 0; x < nqp; ++x) {
    const int32_t node = d.cell_nodes[x];
    MFG_REQUIRE(node >= 0 && node < n_nodes, MFG_ERR_ARGUMENT, "cell node index out of range");
  }
  colour_cells(d, N3);
  // device layout in colour-major cell order
  std::vector<uint32_t> map((size_t)nqp);
  std::vector<double> vert((size_t)n_cells * 24);
  std::vector<char> touched((size_t)n_nodes, 0);
  for (int c = 0; c < n_cells; ++c) {
    const int64_t o = perm[c];
    for (int x = 0; x < N3; ++x) {
      const int32_t node = d.cell_nodes[o * N3 + x];
      // bit 31: first (cell, node) pair in launch order, i.e. the first
      // colour that writes this node stores instead of adding
      const uint32_t first = touched[node] ? 0u : (1u << 31);
      touched[node] = 1;
      map[(size_t)c * N3 + x] = (uint32_t)node | ((uint32_t)cmask[node] << kConstraintShift) | first;
    }
    for (int x = 0; x < 24; ++x) vert[(size_t)c * 24 + x] = d.cell_vertices[o * 24 + x];
  }
  // vmult writes every dof from the cells that reference its node (the first
  // colour stores, the GATHER sum starts from the first contribution): a node
  // no cell references would keep stale dst, so such a description is refused
  for (int64_t node = 0; node < n_nodes; ++node)
    MFG_REQUIRE(touched[node], MFG_ERR_ARGUMENT, "node referenced by no cell");
  // affine cells: the mixed terms of the trilinear map vanish (vertex v at
  // (v & 1, v >> 1 & 1, v >> 2 & 1)); then det J_ref is one number per cell
  std::vector<double> det_ref((size_t)n_cells), jac((size_t)n_cells * 18);
  affine_ref = true;
  for (int c = 0; c < n_cells; ++c) {
    const double *X = &vert[(size_t)c * 24];
    auto V = [&](int v, int e) { return X[v * 3 + e]; };
    double J[3][3], h = 0, mixed = 0;
    for (int e = 0; e < 3; ++e) {
      J[e][0] = V(1, e) - V(0, e);
      J[e][1] = V(2, e) - V(0, e);
      J[e][2] = V(4, e) - V(0, e);
      for (int f = 0; f < 3; ++f) h = std::max(h, std::fabs(J[e][f]));
      mixed = std::max({mixed, std::fabs(V(3, e) - V(2, e) - V(1, e) + V(0, e)),
                        std::fabs(V(5, e) - V(4, e) - V(1, e) + V(0, e)),
                        std::fabs(V(6, e) - V(4, e) - V(2, e) + V(0, e)),
                        std::fabs(V(7, e) - V(6, e) - V(5, e) + V(4, e) - V(3, e) + V(2, e) + V(1, e) - V(0, e))});
    }
    if (mixed > 1e-12 * h) affine_ref = false;
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    det_ref[c] = std::fabs(det);
    // J_ref and J_ref^{-1} of an affine cell (the tensor2 gradient storage)
    double *G = &jac[(size_t)c * 18];
    for (int f = 0; f < 9; ++f) G[f] = J[f / 3][f % 3];
    const double id = det != 0 ? 1.0 / det : 0.0;
    G[9] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    G[10] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    G[11] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    G[12] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    G[13] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    G[14] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    G[15] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    G[16] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    G[17] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  }
  // axis-aligned cells: J_ref and J_ref^{-1} exactly zero off the diagonal
  cartesian_ref = affine_ref;
  for (size_t x = 0; x < jac.size() && cartesian_ref; ++x)
    if ((x % 9) % 4 != 0 && jac[x] != 0.0) cartesian_ref = false;
  upload(cell_nodes, map.data(), map.size(), stream);
  upload(constrained, d.constrained_dofs, (size_t)n_constrained, stream);
  host_constrained.assign(d.constrained_dofs, d.constrained_dofs + n_constrained);
  upload(vertices, vert.data(), vert.size(), stream);
  upload(perm_dev, perm.data(), perm.size(), stream);
  host_cell_nodes.assign(d.cell_nodes, d.cell_nodes + nqp);
  host_cmask = std::move(cmask);
  flag.allocate(sizeof(int));
  const Tables1D t = make_tables(p);
  if (precision == MFG_FP64) upload_tables<double>(*this, t);
  else upload_tables<float>(*this, t);
  // small operators: their colour launches are bound by one workgroup's
  // critical path each, so the two launches of GATHER are faster (measured on
  // the beam's MG levels up to 256K points; equal at 2M, slower at 16M)
  if (nqp <= kGatherMaxPoints) {
    build_gather();
    cell_schedule = MFG_CELLS_GATHER;
  }
  if (affine_ref) {
    const int N = p + 1;
    std::vector<double> w3(N3);
    for (int x = 0; x < N3; ++x) w3[x] = t.weight[x % N] * t.weight[(x / N) % N] * t.weight[x / (N * N)];
    if (precision == MFG_FP64) {
      upload(cell_jxw, det_ref.data(), det_ref.size(), stream);
      upload(wq3, w3.data(), w3.size(), stream);
      if (kind == 0) upload(cell_jac, jac.data(), jac.size(), stream);
    } else {
      std::vector<float> a(det_ref.begin(), det_ref.end()), b(w3.begin(), w3.end());
      upload(cell_jxw, a.data(), a.size(), stream);
      upload(wq3, b.data(), b.size(), stream);
      if (kind == 0) {
        std::vector<float> g(jac.begin(), jac.end());
        upload(cell_jac, g.data(), g.size(), stream);
      }
    }
  }
}

template <typename Number>
static void heat_geometry(OperatorData &op, const mfg_level_desc &d, const double *coef) {
  // per (cell, q): Jinv (9), JxW, c -- MappingQ1 data (heat_transfer.h:483-496)
  const Tables1D t = make_tables(op.p);
  const int N = op.p + 1, N3 = N * N * N;
  const int64_t n = op.nqp;
  if (op.affine_ref) {
    // affine cells: the geometry once per cell, the coefficient per point
    std::vector<Number> geo((size_t)op.n_cells * 10), cq((size_t)n);
    for (int c = 0; c < op.n_cells; ++c) {
      const int64_t orig = op.perm[c];
      const double *X = d.cell_vertices + orig * 24;
      double J[3][3];
      for (int e = 0; e < 3; ++e) {
        J[e][0] = X[3 + e] - X[e];
        J[e][1] = X[6 + e] - X[e];
        J[e][2] = X[12 + e] - X[e];
      }
      const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                         J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      MFG_REQUIRE(det > 0, MFG_ERR_ARGUMENT, "inverted cell");
      const double id = 1.0 / det;
      const double R[9] = {(J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id,
                           (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id, (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id,
                           (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id,
                           (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id,
                           (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id};
      for (int f = 0; f < 9; ++f) geo[(size_t)c * 10 + f] = (Number)R[f];
      geo[(size_t)c * 10 + 9] = (Number)det;
      for (int q = 0; q < N3; ++q) cq[(size_t)c * N3 + q] = (Number)(coef ? coef[orig * N3 + q] : 1.0);
    }
    upload(op.cell_geo, geo.data(), geo.size(), op.stream);
    upload(op.qd, cq.data(), cq.size(), op.stream);
    return;
  }
  std::vector<Number> h((size_t)11 * op.nqp);
  for (int c = 0; c < op.n_cells; ++c)
    for (int q = 0; q < N3; ++q) {
      const int64_t orig = op.perm[c];  // colour-major cell c is input cell orig
      const int i = q % N, j = (q / N) % N, k = q / (N * N);
      const double x[3] = {t.points[i], t.points[j], t.points[k]};
      double J[3][3] = {};
      for (int v = 0; v < 8; ++v) {
        const int b[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
        double nn[3], dd[3];
        for (int e = 0; e < 3; ++e) {
          nn[e] = b[e] ? x[e] : 1 - x[e];
          dd[e] = b[e] ? 1 : -1;
        }
        const double g[3] = {dd[0] * nn[1] * nn[2], nn[0] * dd[1] * nn[2], nn[0] * nn[1] * dd[2]};
        for (int e = 0; e < 3; ++e)
          for (int f = 0; f < 3; ++f) J[e][f] += d.cell_vertices[orig * 24 + v * 3 + e] * g[f];
      }
      const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      MFG_REQUIRE(det > 0, MFG_ERR_ARGUMENT, "inverted cell");
      const double id = 1.0 / det;
      const double R[9] = {(J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id,
                           (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id, (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id,
                           (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id,
                           (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id,
                           (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id};
      const int64_t o = (int64_t)c * N3 + q;
      for (int f = 0; f < 9; ++f) h[f * n + o] = (Number)R[f];
      h[9 * n + o] = (Number)(det * t.weight[i] * t.weight[j] * t.weight[k]);
      h[10 * n + o] = (Number)(coef ? coef[orig * N3 + q] : 1.0);
    }
  upload(op.qd, h.data(), h.size(), op.stream);
}

void OperatorData::setup_heat(const mfg_level_desc &d, const double *coef) {
  if (precision == MFG_FP64) heat_geometry<double>(*this, d, coef);
  else heat_geometry<float>(*this, d, coef);
  n_fields = affine_ref ? 1 : QHeat<double>::kFields;
  ready = true;
}

void OperatorData::setup_solid() {
  // material.h:78-94
  auto constants = [&](double mu, double nu, double rho) {
    MaterialConstants M;
    M.mu = mu;
    M.kappa = (2.0 * mu * (1.0 + nu)) / (3.0 * (1.0 - 2.0 * nu));
    M.c_1 = mu / 2.0;
    M.lambda = (2.0 * mu * nu) / (1.0 - 2.0 * nu);
    M.rho = rho;
    M.rho_alpha = rho * sp.alpha_1;
    return M;
  };
  mats[0] = constants(sp.mu, sp.nu, sp.rho);
  mats[1] = sp.inclusion_mu != 0 ? constants(sp.inclusion_mu, sp.inclusion_nu, sp.inclusion_rho) : mats[0];
  kappa = mats[0].kappa;
  c_1 = mats[0].c_1;
  lambda = mats[0].lambda;
  rho_alpha = mats[0].rho_alpha;
  switch (sp.caching) {
    case MFG_CACHING_SCALAR_REFERENTIAL: n_fields = QSolidSR<double>::kFields; break;
    case MFG_CACHING_TENSOR4: n_fields = QSolidT4<double>::kFields; break;
    case MFG_CACHING_TENSOR4_NS: n_fields = QSolidT4NS<double>::kFields; break;
    default:
      n_fields = sp.formulation == 0 ? QSolidT2F0<double>::kFields
                 : t2_gradient()     ? QSolidT2F1G<double>::kFields
                 : affine_ref        ? QSolidT2F1<double>::kFields - 1
                                     : QSolidT2F1<double>::kFields;
  }
  const size_t es = precision == MFG_FP64 ? 8 : 4;
  qd.allocate((size_t)n_fields * nqp * es);
}

size_t OperatorData::memory_consumption() const {
  return qd.bytes + cell_nodes.bytes + constrained.bytes + u_nodal.bytes + cell_jxw.bytes + wq3.bytes + cell_geo.bytes + cell_jac.bytes +
         gather_offsets.bytes + gather_points.bytes + gather_cmask.bytes + contrib.bytes;
}

}  // namespace mfg

// --------------------------------------------------------------- C ABI ---
extern "C" {

int mfg_solid_create(const mfg_level_desc *level, const mfg_solid_params *params, int32_t precision,
                     mfg_operator *out) {
  return guarded([&] {
    MFG_REQUIRE(level && params && out, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(precision == MFG_FP64 || precision == MFG_FP32, MFG_ERR_ARGUMENT, "bad precision");
    MFG_REQUIRE(params->formulation == 0 || params->formulation == 1, MFG_ERR_ARGUMENT,
                "Unknown material formulation");
    MFG_REQUIRE(params->caching >= MFG_CACHING_SCALAR_REFERENTIAL && params->caching <= MFG_CACHING_TENSOR4_NS,
                MFG_ERR_ARGUMENT, "Unknown caching strategy");
    // mf_nh_operator.h:706: formulation 0 is implemented for tensor2 only
    MFG_REQUIRE(params->formulation == 1 || params->caching == MFG_CACHING_TENSOR2, MFG_ERR_ARGUMENT,
                "formulation 0 needs the tensor2 caching");
    MFG_REQUIRE(params->mu > 0 && params->nu > -1 && params->nu < 0.5, MFG_ERR_ARGUMENT, "bad material");
    MFG_REQUIRE(params->inclusion_mu == 0 ||
                    (params->inclusion_mu > 0 && params->inclusion_nu > -1 && params->inclusion_nu < 0.5),
                MFG_ERR_ARGUMENT, "bad inclusion material");
    auto op = new mfg_operator_t();
    try {
      op->kind = 0;
      op->precision = precision;
      op->sp = *params;
      op->setup_layout(*level);
      // tensor2 formulation 1 on all-affine cells: the 10-word point storage
      // (faster and smaller than the folded 16 words; the same operator)
      if (params->formulation == 1 && params->caching == MFG_CACHING_TENSOR2 && op->affine_ref)
        op->t2_storage = MFG_T2_GRADIENT;
      op->setup_solid();
    } catch (...) {
      delete op;
      throw;
    }
    *out = op;
  });
}

int mfg_heat_create(const mfg_level_desc *level, double dt, const double *coefficient, int32_t precision,
                    mfg_operator *out) {
  return guarded([&] {
    MFG_REQUIRE(level && out, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(dt > 0, MFG_ERR_ARGUMENT, "delta_t must be positive");
    MFG_REQUIRE(precision == MFG_FP64 || precision == MFG_FP32, MFG_ERR_ARGUMENT, "bad precision");
    auto op = new mfg_operator_t();
    try {
      op->kind = 1;
      op->precision = precision;
      op->dt = dt;
      op->setup_layout(*level);
      op->setup_heat(*level, coefficient);
    } catch (...) {
      delete op;
      throw;
    }
    *out = op;
  });
}

int mfg_operator_destroy(mfg_operator op) {
  return guarded([&] { delete op; });
}

int mfg_operator_set_stream(mfg_operator op, void *s) {
  return guarded([&] {
    MFG_REQUIRE(op, MFG_ERR_ARGUMENT, "null operator");
    op->stream = (hipStream_t)s;
  });
}

int64_t mfg_operator_m(mfg_operator op) { return op ? op->n_dofs : -1; }
int64_t mfg_operator_n(mfg_operator op) { return op ? op->n_dofs : -1; }
int32_t mfg_operator_precision(mfg_operator op) { return op ? op->precision : -1; }

int mfg_solid_cache(mfg_operator op, const void *u, int32_t u_precision) {
  return guarded([&] {
    MFG_REQUIRE(op, MFG_ERR_ARGUMENT, "null operator");
    op->cache(u, u_precision);
  });
}

int mfg_operator_vmult(mfg_operator op, void *dst, const void *src) {
  return guarded([&] {
    MFG_REQUIRE(op, MFG_ERR_ARGUMENT, "null operator");
    op->apply(src, dst, false);
  });
}
int mfg_operator_vmult_add(mfg_operator op, void *dst, const void *src) {
  return guarded([&] {
    MFG_REQUIRE(op, MFG_ERR_ARGUMENT, "null operator");
    op->apply(src, dst, true);
  });
}
int mfg_operator_Tvmult(mfg_operator op, void *dst, const void *src) { return mfg_operator_vmult(op, dst, src); }
int mfg_operator_Tvmult_add(mfg_operator op, void *dst, const void *src) {
  return mfg_operator_vmult_add(op, dst, src);
}

int mfg_operator_compute_diagonal(mfg_operator op) {
  return guarded([&] {
    MFG_REQUIRE(op, MFG_ERR_ARGUMENT, "null operator");
    op->compute_diagonal();
  });
}

int mfg_operator_copy_diagonal(mfg_operator op, void *diag, void *inv_diag) {
  return guarded([&] {
    MFG_REQUIRE(op, MFG_ERR_ARGUMENT, "null operator");
    MFG_REQUIRE(op->has_diagonal, MFG_ERR_STATE, "compute_diagonal() not called");
    if (diag) MFG_HIP(hipMemcpyAsync(diag, op->diag.ptr, op->diag.bytes, hipMemcpyDeviceToDevice, op->stream));
    if (inv_diag)
      MFG_HIP(hipMemcpyAsync(inv_diag, op->inv_diag.ptr, op->inv_diag.bytes, hipMemcpyDeviceToDevice, op->stream));
  });
}

int mfg_operator_el(mfg_operator op, int64_t row, int64_t col, double *value) {
  return guarded([&] {
    MFG_REQUIRE(op && value, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(row == col, MFG_ERR_ARGUMENT, "el() only provides diagonal entries");
    MFG_REQUIRE(row >= 0 && row < op->n_dofs, MFG_ERR_ARGUMENT, "row out of range");
    MFG_REQUIRE(op->has_diagonal, MFG_ERR_STATE, "compute_diagonal() not called");
    if (op->precision == MFG_FP64) {
      MFG_HIP(hipMemcpyAsync(value, op->diag.as<double>() + row, 8, hipMemcpyDeviceToHost, op->stream));
      MFG_HIP(hipStreamSynchronize(op->stream));
    } else {
      float f = 0;
      MFG_HIP(hipMemcpyAsync(&f, op->diag.as<float>() + row, 4, hipMemcpyDeviceToHost, op->stream));
      MFG_HIP(hipStreamSynchronize(op->stream));
      *value = f;
    }
  });
}

int mfg_operator_precondition_jacobi(mfg_operator op, void *dst, const void *src, double omega) {
  return guarded([&] {
    MFG_REQUIRE(op && dst && src, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(op->has_diagonal, MFG_ERR_STATE, "compute_diagonal() not called");
    const int threads = 256, blocks = std::min(blocks_for(op->n_dofs, threads), 65536);
    if (op->precision == MFG_FP64)
      scaled_product_kernel<double><<<blocks, threads, 0, op->stream>>>((double *)dst, op->inv_diag.as<double>(),
                                                                        (const double *)src, omega, op->n_dofs);
    else
      scaled_product_kernel<float><<<blocks, threads, 0, op->stream>>>((float *)dst, op->inv_diag.as<float>(),
                                                                       (const float *)src, (float)omega, op->n_dofs);
    MFG_HIP(hipGetLastError());
  });
}

size_t mfg_operator_memory_consumption(mfg_operator op) { return op ? op->memory_consumption() : 0; }

int mfg_operator_enable_timing(mfg_operator op, int32_t enable) {
  return guarded([&] {
    MFG_REQUIRE(op, MFG_ERR_ARGUMENT, "null operator");
    op->timing = enable != 0;
    op->events_used = 0;
  });
}

int mfg_operator_timing(mfg_operator op, double *total_ms, int64_t *launches) {
  return guarded([&] {
    MFG_REQUIRE(op && total_ms && launches, MFG_ERR_ARGUMENT, "null argument");
    op->timing_collect(*total_ms, *launches);
  });
}

int mfg_operator_qp_data(mfg_operator op, void *out, int32_t *n_fields) {
  return guarded([&] {
    MFG_REQUIRE(op, MFG_ERR_ARGUMENT, "null operator");
    if (n_fields) *n_fields = op->n_fields;
    if (out) {
      MFG_REQUIRE(op->ready, MFG_ERR_STATE, "no quadrature data yet");
      const int n3 = (op->p + 1) * (op->p + 1) * (op->p + 1);
      const int threads = 256;
      const int blocks = std::min(blocks_for(op->nqp * op->n_fields, threads), 65536);
      if (op->precision == MFG_FP64)
        unpermute_fields_kernel<double><<<blocks, threads, 0, op->stream>>>(
            op->qd.as<double>(), (double *)out, op->perm_dev.as<int32_t>(), op->n_cells, n3, op->n_fields,
            op->kind == 0 && op->sp.caching == MFG_CACHING_TENSOR2 && op->sp.formulation == 1 ? 1 : 0);
      else
        unpermute_fields_kernel<float><<<blocks, threads, 0, op->stream>>>(
            op->qd.as<float>(), (float *)out, op->perm_dev.as<int32_t>(), op->n_cells, n3, op->n_fields,
            op->kind == 0 && op->sp.caching == MFG_CACHING_TENSOR2 && op->sp.formulation == 1 ? 1 : 0);
      MFG_HIP(hipGetLastError());
    }
  });
}

int32_t mfg_operator_n_colours(mfg_operator op) { return op ? (int32_t)op->colour_start.size() - 1 : -1; }

int mfg_operator_set_cell_schedule(mfg_operator op, int32_t schedule) {
  return guarded([&] {
    MFG_REQUIRE(op, MFG_ERR_ARGUMENT, "null operator");
    MFG_REQUIRE(schedule == MFG_CELLS_COLOURS || schedule == MFG_CELLS_GATHER, MFG_ERR_ARGUMENT,
                "unknown cell schedule");
    op->cell_schedule = MFG_CELLS_COLOURS;
    if (schedule == MFG_CELLS_GATHER) {
      if (!op->contrib.ptr) op->build_gather();
      op->cell_schedule = schedule;
    }
    ++op->buffers_version;  // another kernel path: captured graphs are stale
  });
}

int32_t mfg_operator_cell_schedule(mfg_operator op) { return op ? op->cell_schedule : -1; }

int mfg_operator_set_point_storage(mfg_operator op, int32_t storage) {
  return guarded([&] {
    MFG_REQUIRE(op, MFG_ERR_ARGUMENT, "null operator");
    MFG_REQUIRE(storage == MFG_T2_FOLDED || storage == MFG_T2_GRADIENT, MFG_ERR_ARGUMENT, "unknown point storage");
    if (storage == op->t2_storage) return;
    MFG_REQUIRE(op->kind == 0 && op->sp.caching == MFG_CACHING_TENSOR2 && op->sp.formulation == 1, MFG_ERR_ARGUMENT,
                "the point storage applies to the tensor2 caching of formulation 1");
    MFG_REQUIRE(storage == MFG_T2_FOLDED || op->affine_ref, MFG_ERR_ARGUMENT,
                "the gradient storage needs a mesh whose cells are all affine");
    MFG_HIP(hipStreamSynchronize(op->stream));
    op->t2_storage = storage;
    op->setup_solid();  // reallocates the point cache for the new field count
    op->ready = false;
    op->has_diagonal = false;
    ++op->buffers_version;
  });
}

int32_t mfg_operator_point_storage(mfg_operator op) { return op ? op->t2_storage : -1; }

}  // extern "C"
