// C++ host mirror of the reference's operator interfaces on top of the
// mfgpu C ABI (include/mfgpu.h).  Method names, argument meaning and error
// behaviour follow the reference classes so that solver code written against
// them compiles unchanged:
//   NeoHookOperator<dim,Number>  include/solid_mechanics/mf_nh_operator.h:99-237
//   LaplaceOperator<dim,number>  include/heat_transfer/heat_transfer.h:81-127
//   SolverCG / PreconditionMG / PreconditionJacobi as used at
//   mf_elasticity.h:1981-2070.
// Non-zero statuses become mfgpu::ExcMfg exceptions (the AssertThrow of the
// reference).  Vectors are device vectors (mfgpu::DeviceVector) in the
// operator's precision.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mfgpu.h"

namespace mfgpu {

struct ExcMfg : std::runtime_error {
  int code;
  ExcMfg(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int status) {
  if (status != MFG_OK) throw ExcMfg(status, mfg_last_error());
}

template <typename Number>
constexpr int32_t precision_of() {
  return sizeof(Number) == 8 ? MFG_FP64 : MFG_FP32;
}

// Minimal device vector (LinearAlgebra::distributed::Vector collapsed to one GPU).
template <typename Number>
class DeviceVector {
 public:
  DeviceVector() = default;
  explicit DeviceVector(int64_t n) { reinit(n); }
  DeviceVector(const DeviceVector &) = delete;
  DeviceVector &operator=(const DeviceVector &) = delete;
  ~DeviceVector() { release(); }
  void reinit(int64_t n) {
    release();
    n_ = n;
    if (n > 0 && hipMalloc(&ptr_, n * sizeof(Number)) != hipSuccess) throw ExcMfg(MFG_ERR_HIP, "hipMalloc failed");
    *this = Number(0);
  }
  DeviceVector &operator=(Number v) {
    std::vector<Number> h(n_, v);
    copy_from_host(h);
    return *this;
  }
  void copy_from_host(const std::vector<Number> &h) {
    if ((int64_t)h.size() != n_) throw ExcMfg(MFG_ERR_ARGUMENT, "size mismatch");
    if (n_ && hipMemcpy(ptr_, h.data(), n_ * sizeof(Number), hipMemcpyHostToDevice) != hipSuccess)
      throw ExcMfg(MFG_ERR_HIP, "hipMemcpy failed");
  }
  std::vector<Number> to_host() const {
    std::vector<Number> h(n_);
    if (n_ && hipMemcpy(h.data(), ptr_, n_ * sizeof(Number), hipMemcpyDeviceToHost) != hipSuccess)
      throw ExcMfg(MFG_ERR_HIP, "hipMemcpy failed");
    return h;
  }
  int64_t size() const { return n_; }
  Number *data() { return ptr_; }
  const Number *data() const { return ptr_; }

 private:
  void release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    n_ = 0;
  }
  Number *ptr_ = nullptr;
  int64_t n_ = 0;
};

// Structured box level (subdivided_hyper_rectangle + distribute_dofs + constraints).
class BoxLevel {
 public:
  BoxLevel(const double lo[3], const double hi[3], const int32_t n[3], int degree, int n_components, int face_mask,
           int comp_mask, int numbering = MFG_NUMBERING_LEXICOGRAPHIC) {
    check(mfg_box_mesh_create(lo, hi, n, degree, n_components, face_mask, comp_mask, &mesh_));
    check(mfg_box_mesh_renumber(mesh_, numbering));
    check(mfg_box_mesh_describe(mesh_, &desc_));
  }
  ~BoxLevel() { mfg_box_mesh_destroy(mesh_); }
  BoxLevel(const BoxLevel &) = delete;
  const mfg_level_desc &desc() const { return desc_; }
  mfg_mesh handle() const { return mesh_; }
  int64_t n_dofs() const { return desc_.n_nodes * desc_.n_components; }

 private:
  mfg_mesh mesh_ = nullptr;
  mfg_level_desc desc_{};
};

// Common matrix interface (vmult, Tvmult, m, n, el, diagonal).
template <typename Number>
class OperatorBase {
 public:
  using value_type = Number;
  using VectorType = DeviceVector<Number>;
  ~OperatorBase() { clear(); }
  void clear() {
    if (op_) mfg_operator_destroy(op_);
    op_ = nullptr;
  }
  unsigned int m() const { return (unsigned int)mfg_operator_m(op_); }
  unsigned int n() const { return (unsigned int)mfg_operator_n(op_); }
  void vmult(VectorType &dst, const VectorType &src) const { check(mfg_operator_vmult(op_, dst.data(), src.data())); }
  void Tvmult(VectorType &dst, const VectorType &src) const { check(mfg_operator_Tvmult(op_, dst.data(), src.data())); }
  void vmult_add(VectorType &dst, const VectorType &src) const {
    check(mfg_operator_vmult_add(op_, dst.data(), src.data()));
  }
  void Tvmult_add(VectorType &dst, const VectorType &src) const {
    check(mfg_operator_Tvmult_add(op_, dst.data(), src.data()));
  }
  void compute_diagonal() { check(mfg_operator_compute_diagonal(op_)); }
  Number el(const unsigned int row, const unsigned int col) const {
    double v = 0;
    check(mfg_operator_el(op_, row, col, &v));
    return (Number)v;
  }
  void precondition_Jacobi(VectorType &dst, const VectorType &src, const Number omega) const {
    check(mfg_operator_precondition_jacobi(op_, dst.data(), src.data(), omega));
  }
  std::size_t memory_consumption() const { return mfg_operator_memory_consumption(op_); }
  mfg_operator handle() const { return op_; }

 protected:
  mfg_operator op_ = nullptr;
};

// NeoHookOperator<dim, Number>: initialize + set_material + cache + vmult.
template <int dim, typename Number>
class NeoHookOperator : public OperatorBase<Number> {
  static_assert(dim == 3, "the device operator is three-dimensional");

 public:
  struct Material {  // Material_Compressible_Neo_Hook_One_Field constructor arguments
    double mu, nu, rho, alpha_1;
    unsigned int formulation;
  };
  void initialize(const mfg_level_desc &level, const std::string &caching, const Material &mat) {
    this->clear();
    // the "MF caching" names of the reference's parameter file (mf_nh_operator.h:346-387)
    if (caching == "scalar_referential") caching_ = MFG_CACHING_SCALAR_REFERENTIAL;
    else if (caching == "tensor2") caching_ = MFG_CACHING_TENSOR2;
    else if (caching == "tensor4") caching_ = MFG_CACHING_TENSOR4;
    else if (caching == "tensor4_ns") caching_ = MFG_CACHING_TENSOR4_NS;
    else throw ExcMfg(MFG_ERR_ARGUMENT, "Unknown caching");
    level_ = level;
    create(mat, nullptr);
  }
  // set_material(material, material_inclusion) (mf_nh_operator.h:529-539):
  // the inclusion material acts on the cells whose material id (level's
  // material_ids) is not 0.  Rebuilds the operator; call cache() afterwards.
  void set_material(const Material &mat, const Material &inclusion) { create(mat, &inclusion); }

 private:
  void create(const Material &mat, const Material *inclusion) {
    this->clear();
    mfg_solid_params p{mat.mu, mat.nu, mat.rho, mat.alpha_1, (int32_t)mat.formulation, caching_, 0, 0, 0};
    if (inclusion) {
      p.inclusion_mu = inclusion->mu;
      p.inclusion_nu = inclusion->nu;
      p.inclusion_rho = inclusion->rho;
    }
    check(mfg_solid_create(&level_, &p, precision_of<Number>(), &this->op_));
  }
  mfg_level_desc level_{};
  int32_t caching_ = MFG_CACHING_TENSOR2;

 public:
  template <typename UNumber>
  void cache(const DeviceVector<UNumber> &displacement) {
    check(mfg_solid_cache(this->op_, displacement.data(), precision_of<UNumber>()));
  }
  // MFG_T2_FOLDED or MFG_T2_GRADIENT (tensor2, formulation 1); call cache() again afterwards
  void set_point_storage(const int32_t storage) { check(mfg_operator_set_point_storage(this->op_, storage)); }
};

// LaplaceOperator<dim, number>: M + dt c K.
template <int dim, typename number>
class LaplaceOperator : public OperatorBase<number> {
 public:
  void initialize(const mfg_level_desc &level, const std::vector<double> &coefficient = {}) {
    this->clear();
    level_ = level;
    coef_ = coefficient;
  }
  void set_delta_t(const double dt) {
    this->clear();
    check(mfg_heat_create(&level_, dt, coef_.empty() ? nullptr : coef_.data(), precision_of<number>(), &this->op_));
  }

 private:
  mfg_level_desc level_{};
  std::vector<double> coef_;
};

// PreconditionIdentity / PreconditionJacobi / PreconditionMG for SolverCG.
struct PreconditionIdentity {};
template <typename Op>
struct PreconditionJacobi {
  const Op *op = nullptr;
  double omega = 1.0;
  void initialize(const Op &o, double w) {
    op = &o;
    omega = w;
  }
};
class PreconditionMG {
 public:
  PreconditionMG(std::vector<mfg_operator> ops, std::vector<mfg_transfer> transfers, const mfg_mg_params &prm)
      : ops_(std::move(ops)), transfers_(std::move(transfers)) {
    check(mfg_mg_create((int32_t)ops_.size(), ops_.data(), transfers_.data(), &prm, &mg_));
    check(mfg_mg_setup(mg_));
  }
  ~PreconditionMG() { mfg_mg_destroy(mg_); }
  PreconditionMG(const PreconditionMG &) = delete;
  mfg_mg handle() const { return mg_; }

 private:
  std::vector<mfg_operator> ops_;
  std::vector<mfg_transfer> transfers_;
  mfg_mg mg_ = nullptr;
};

// SolverControl + SolverCG (relative tolerance, iteration cap; NoConvergence
// surfaces as ExcMfg with code MFG_ERR_NOCONV).
struct SolverControl {
  int max_steps = 1000;
  double tolerance = 1e-8;  // relative to ||b||
  mfg_solve_info info{};
  unsigned int last_step() const { return (unsigned int)info.iterations; }
  double last_value() const { return info.final_residual; }
};

template <typename VectorType>
class SolverCG {
 public:
  explicit SolverCG(SolverControl &c) : control_(c) {}
  template <typename Op>
  void solve(const Op &A, VectorType &x, const VectorType &b, const PreconditionIdentity &) {
    check(mfg_cg_solve(A.handle(), MFG_PRECOND_NONE, nullptr, 1.0, b.data(), x.data(), control_.tolerance,
                       control_.max_steps, &control_.info));
  }
  template <typename Op>
  void solve(const Op &A, VectorType &x, const VectorType &b, const PreconditionJacobi<Op> &P) {
    check(mfg_cg_solve(A.handle(), MFG_PRECOND_JACOBI, nullptr, P.omega, b.data(), x.data(), control_.tolerance,
                       control_.max_steps, &control_.info));
  }
  template <typename Op>
  void solve(const Op &A, VectorType &x, const VectorType &b, const PreconditionMG &P) {
    check(mfg_cg_solve(A.handle(), MFG_PRECOND_MG, P.handle(), 1.0, b.data(), x.data(), control_.tolerance,
                       control_.max_steps, &control_.info));
  }

 private:
  SolverControl &control_;
};

}  // namespace mfgpu
