// Drop-in check of the C++ host mirror (include/mfgpu.hpp): the solver code
// below is written the way Solid::solve_linear_system uses NeoHookOperator,
// PreconditionMG and SolverCG (mf_elasticity.h:1981-2070), on the small
// configuration (unit cube, 8^3 Q2 cells, 3 levels, Chebyshev degree 4).
// Prints "DROPIN OK iterations=<k> error=<e>" on success, exits non-zero
// otherwise.
#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

#include "../../include/mfgpu.hpp"

int main() {
  using namespace mfgpu;
  try {
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    const double E = 1e6, nu = 0.3, rho = 1000, beta = 0.25, dt = 1e-3;
    const NeoHookOperator<3, double>::Material mat{E / (2 * (1 + nu)), nu, rho, 1.0 / (beta * dt * dt), 1};
    std::vector<std::unique_ptr<BoxLevel>> levels;
    std::vector<std::unique_ptr<NeoHookOperator<3, double>>> ops;
    for (int l = 0; l < 3; ++l) {
      const int32_t n[3] = {2 << l, 2 << l, 2 << l};
      levels.emplace_back(new BoxLevel(lo, hi, n, 2, 3, 1, 7));
      ops.emplace_back(new NeoHookOperator<3, double>());
      ops.back()->initialize(levels.back()->desc(), "tensor2", mat);
    }
    // transfers and interpolate_to_mg of the linearisation point
    std::vector<std::vector<int32_t>> children(2);
    std::vector<mfg_transfer> transfers(2);
    for (int l = 1; l < 3; ++l) {
      children[l - 1].resize(levels[l - 1]->desc().n_cells * 8);
      check(mfg_box_mesh_children(levels[l - 1]->handle(), levels[l]->handle(), children[l - 1].data()));
      check(mfg_transfer_create(ops[l - 1]->handle(), ops[l]->handle(), children[l - 1].data(), &transfers[l - 1]));
    }
    std::vector<std::unique_ptr<DeviceVector<double>>> u(3);
    for (int l = 0; l < 3; ++l) u[l].reset(new DeviceVector<double>(levels[l]->n_dofs()));
    std::vector<double> uh(levels[2]->n_dofs());
    check(mfg_box_mesh_fill_u0(levels[2]->handle(), 1e-3, 1e-5, 0, uh.data()));
    u[2]->copy_from_host(uh);
    for (int l = 2; l >= 1; --l) check(mfg_transfer_interpolate(transfers[l - 1], u[l - 1]->data(), u[l]->data()));
    for (int l = 0; l < 3; ++l) {
      ops[l]->cache(*u[l]);  // setup_operator_cache: cache() + compute_diagonal()
      ops[l]->compute_diagonal();
    }
    const mfg_mg_params prm{4, 20.0, 30, 0, 1e-3, 1e-3};
    PreconditionMG mg({ops[0]->handle(), ops[1]->handle(), ops[2]->handle()}, transfers, prm);

    const NeoHookOperator<3, double> &A = *ops[2];
    const int64_t n = levels[2]->n_dofs();
    DeviceVector<double> x_true(n), b(n), x(n);
    check(mfg_fill_uniform(2, n, MFG_FP64, x_true.data(), nullptr));
    std::vector<double> xt = x_true.to_host();
    for (int64_t i = 0; i < levels[2]->desc().n_constrained; ++i) xt[levels[2]->desc().constrained_dofs[i]] = 0.0;
    x_true.copy_from_host(xt);
    A.vmult(b, x_true);

    SolverControl control;
    control.max_steps = 100;
    control.tolerance = 1e-8;
    SolverCG<DeviceVector<double>> cg(control);
    x = 0.0;
    cg.solve(A, x, b, mg);
    const std::vector<double> xs = x.to_host();
    double e2 = 0, n2 = 0;
    for (int64_t i = 0; i < n; ++i) {
      e2 += (xs[i] - xt[i]) * (xs[i] - xt[i]);
      n2 += xt[i] * xt[i];
    }
    const double err = std::sqrt(e2 / n2);
    // Tvmult == vmult, el() == diagonal, PreconditionJacobi path
    DeviceVector<double> y1(n), y2(n);
    A.vmult(y1, x_true);
    A.Tvmult(y2, x_true);
    const std::vector<double> h1 = y1.to_host(), h2 = y2.to_host();
    for (int64_t i = 0; i < n; ++i)
      if (h1[i] != h2[i]) throw ExcMfg(MFG_ERR_NUMERIC, "Tvmult differs from vmult");
    if (!(A.el(7, 7) > 0)) throw ExcMfg(MFG_ERR_NUMERIC, "non-positive diagonal");
    SolverControl jc;
    jc.max_steps = 2000;
    SolverCG<DeviceVector<double>> cgj(jc);
    PreconditionJacobi<NeoHookOperator<3, double>> jac;
    jac.initialize(A, 0.65);
    x = 0.0;
    cgj.solve(A, x, b, jac);
    for (auto t : transfers) mfg_transfer_destroy(t);
    std::printf("DROPIN OK iterations=%u jacobi_iterations=%u error=%.3e\n", control.last_step(), jc.last_step(), err);
    return (err < 1e-6 && control.last_step() < 20) ? 0 : 1;
  } catch (const ExcMfg &e) {
    std::printf("DROPIN FAILED code=%d %s\n", e.code, e.what());
    return 2;
  }
}
