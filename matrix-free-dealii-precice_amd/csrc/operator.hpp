// Device operator object shared by the operator, multigrid and solver code.
#pragma once
#include "common.hpp"

namespace mfg {

// operators with at most this many quadrature points use the GATHER cell
// schedule unless another is set
constexpr int64_t kGatherMaxPoints = 1 << 20;

// device bytes of the GATHER schedule's tables and contribution buffer
inline size_t gather_bytes(int64_t nqp, int64_t n_nodes, int nc, size_t es) {
  return (size_t)nqp * (nc * es + 4) + (size_t)(n_nodes + 1) * 4 + (size_t)n_nodes;
}

struct OperatorData {
  int kind = 0;  // 0 solid (NeoHookOperator), 1 heat (LaplaceOperator)
  int precision = MFG_FP64;
  int p = 1, nc = 1, n_cells = 0, n_fields = 0;
  int64_t n_nodes = 0, n_dofs = 0, nqp = 0, n_constrained = 0;
  hipStream_t stream = nullptr;
  mfg_solid_params sp{};
  double kappa = 0, c_1 = 0, lambda = 0, rho_alpha = 0, dt = 0;  // material 0
  // neo-Hookean constants (material.h:78-94) of material 0 and of the
  // inclusion material, used on cells whose material id is not 0
  // (mf_nh_operator.h:404, 656-658)
  struct MaterialConstants {
    double mu = 0, lambda = 0, kappa = 0, c_1 = 0, rho = 0, rho_alpha = 0;
  };
  MaterialConstants mats[2];
  bool ready = false, has_diagonal = false;

  DeviceBuffer cell_nodes;   // uint32: node | constraint bits << 28
  DeviceBuffer constrained;  // int32 sorted dof list
  DeviceBuffer vertices;     // double n_cells*24
  DeviceBuffer tab_shape, tab_colloc, tab_sq, tab_points, tab_weights;
  DeviceBuffer qd;           // quadrature cache, field-major SoA
  DeviceBuffer diag, inv_diag;
  DeviceBuffer u_nodal;      // scalar_referential caching: the linearisation point
  // reference-frame JxW per cell: when every cell's trilinear map is affine,
  // JxW_ref = cell_jxw[stored cell] * wq3[point] (tensor2, formulation 1)
  bool affine_ref = false;
  DeviceBuffer cell_jxw, wq3;
  DeviceBuffer cell_geo;  // heat on affine cells: Jinv (9), det J per stored cell
  DeviceBuffer cell_jac;  // solid on affine cells: J_ref (9), J_ref^{-1} (9) per stored cell
  bool cartesian_ref = false;  // affine, and every cell's J_ref (so J_ref^{-1}) is diagonal
  int t2_storage = MFG_T2_FOLDED;  // tensor2 formulation-1 point storage (MFG_T2_*)
  bool t2_gradient() const { return kind == 0 && sp.caching == MFG_CACHING_TENSOR2 && sp.formulation == 1 && t2_storage == MFG_T2_GRADIENT; }
  DeviceBuffer flag;         // int error flag
  std::vector<int32_t> host_constrained;
  std::vector<int32_t> host_cell_nodes;  // input cell order (used to build transfers)
  std::vector<uint8_t> host_cmask;       // per node: constrained components
  // cells are stored colour-major: perm[c] = input index of stored cell c,
  // colour k occupies stored cells [colour_start[k], colour_start[k+1])
  std::vector<int32_t> perm;
  std::vector<int> colour_start;
  // with an inclusion material the cells of colour k are sorted by material
  // (as MatrixFree batches cells by cell_vectorization_category,
  // mf_elasticity.h:837-855): material 0 in [colour_start[k], mat_split[k]),
  // the inclusion in [mat_split[k], colour_start[k+1]).  Empty otherwise.
  std::vector<int> mat_split;
  bool has_inclusion() const { return !mat_split.empty(); }
  // f(colour, first stored cell, end, material) for every non-empty range of
  // one colour and one material, in launch (stored) order
  template <class F>
  void for_each_range(F f) const {
    for (size_t k = 0; k + 1 < colour_start.size(); ++k) {
      const int b = colour_start[k], e = colour_start[k + 1], s = has_inclusion() ? mat_split[k] : e;
      if (s > b) f((int)k, b, s, 0);
      if (e > s) f((int)k, s, e, 1);
    }
  }
  DeviceBuffer perm_dev;
  void colour_cells(const mfg_level_desc &d, int N3);
  int cell_schedule = MFG_CELLS_COLOURS;
  // bumped whenever a device buffer that apply() reads is reallocated (graphs
  // that captured the old addresses must be re-captured)
  uint64_t buffers_version = 0;
  // GATHER schedule: per node, the points (stored cell * N3 + point) that
  // touch it in colour order (CSR), its constrained components, and the
  // per-point contribution buffer
  DeviceBuffer gather_offsets, gather_points, gather_cmask, contrib;
  void build_gather();
  // optional cell-kernel timing
  bool timing = false;
  mutable std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  mutable size_t events_used = 0;
  ~OperatorData();

  void setup_layout(const mfg_level_desc &d);
  void setup_solid();
  void setup_heat(const mfg_level_desc &d, const double *coef);
  void cache(const void *u, int u_precision);
  // dst (+)= A src on `on` (the operator's stream when null)
  void apply(const void *src, void *dst, bool add, hipStream_t on = nullptr) const;
  void compute_diagonal();
  size_t memory_consumption() const;
  size_t elem_size() const { return precision == MFG_FP64 ? 8 : 4; }
  void timing_begin(hipStream_t s) const;
  void timing_end(hipStream_t s) const;
  void timing_collect(double &ms, int64_t &n);
};

}  // namespace mfg

struct mfg_operator_t : mfg::OperatorData {};
