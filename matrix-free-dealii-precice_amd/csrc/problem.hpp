// Device-resident solid problem: level hierarchy, multigrid, Newmark state
// and the Newton-Newmark step (Solid::solve_nonlinear_timestep).
#pragma once
#include <memory>

#include "operator.hpp"
#include "solver.hpp"

namespace mfg {

struct SolidProblem {
  int L = 0;
  int level_precision = MFG_FP64;
  std::vector<std::unique_ptr<mfg_operator_t>> levels;
  std::unique_ptr<mfg_operator_t> fine64;  // outer operator when levels are FP32
  std::vector<std::unique_ptr<mfg_transfer_t>> transfers;
  std::unique_ptr<mfg_mg_t> mg;
  mfg_newmark_params nm{};
  double a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0, a6 = 0;
  DeviceBuffer u, du, delta_u, v, a, v_old, a_old, R;
  // FP32 levels: the condition estimate's right-hand side and iterate in FP32
  DeviceBuffer cond_b, cond_x;
  std::vector<DeviceBuffer> u_lev;
  hipStream_t stream = nullptr;
  CGWorkspace ws;
  Reducer red;
  // coupling interface: (stored cell, face) pairs on the device and the
  // caller's traction per face point, added to every residual
  DeviceBuffer iface_faces;
  int64_t iface_n = 0;
  const double *iface_traction = nullptr;

  OperatorData &fine() { return fine64 ? *fine64 : *levels.back(); }
  int64_t n() { return fine().n_dofs; }
  void set_stream(hipStream_t s);
  void update_matrix_free();
  void step(mfg_step_info &info);
  size_t memory() const;
};

// (input cell, face) pairs -> (stored cell, face) pairs on the device
DeviceBuffer stored_faces(const OperatorData &op, int64_t n_faces, const int32_t *faces);
// rhs += traction term over the faces (stored pairs on the device), stream-ordered
void add_interface_traction(const OperatorData &op, int64_t n_faces, const int32_t *stored, const double *u,
                            const double *traction, double *rhs);

// Builds a problem (levels coarse..fine, children[l-1] between l-1 and l).
std::unique_ptr<SolidProblem> make_solid_problem(int n_levels, const mfg_level_desc *levels,
                                                 const int32_t *const *children, const mfg_solid_params &material,
                                                 const mfg_newmark_params &newmark, const mfg_mg_params &mgp,
                                                 int level_precision, hipStream_t stream);
// Device bytes a problem of these levels will allocate (admission control).
size_t estimate_solid_problem_bytes(int n_levels, const mfg_level_desc *levels, int level_precision);

}  // namespace mfg

struct mfg_solid_problem_t {
  std::unique_ptr<mfg::SolidProblem> p;
};
