// Newton-Newmark time stepping on the device (Solid::solve_nonlinear_timestep
// and friends, mf_elasticity.h:1357-1491, 1600-1620, 1647-1908, 1981-2070).
// Host C++ drives the loop; residual, Newmark updates, operator caches,
// multigrid set-up and the MG-CG solve are device kernels.
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "common.hpp"
#include "operator.hpp"
#include "residual_kernels.hpp"
#include "problem.hpp"
#include "solver.hpp"

namespace mfg {

template <int P>
struct ResidualCPB {
  static constexpr int value = P == 1 ? 8 : P == 2 ? 7 : P == 6 ? 1 : 2;
};

template <int P>
static void residual_launch(const OperatorData &op, const double *u, const double *acc, const double *body,
                            double *R) {
  constexpr int CPB = ResidualCPB<P>::value;
  ResidualArgs<double> a;
  a.cell_nodes = op.cell_nodes.as<uint32_t>();
  a.u = u;
  a.acc = acc;
  a.R = R;
  a.vertices = op.vertices.as<double>();
  a.shape = op.tab_shape.as<double>();
  a.colloc = op.tab_colloc.as<double>();
  a.points = op.tab_points.as<double>();
  a.weights = op.tab_weights.as<double>();
  for (int d = 0; d < 3; ++d) a.body[d] = body[d];
  a.formulation = op.sp.formulation;
  a.error = op.flag.as<int>();
  // one launch per colour and material (the cell's material,
  // mf_elasticity.h:1781-1784)
  op.for_each_range([&](int, int b, int e, int m) {
    const OperatorData::MaterialConstants &M = op.mats[m];
    a.mu = M.mu;
    a.lambda = M.lambda;
    a.kappa = M.kappa;
    a.c_1 = M.c_1;
    a.rho = M.rho;
    a.cell_begin = b;
    a.cell_end = e;
    residual_cells<P, CPB, double><<<blocks_for(e - b, CPB), block_threads<P, CPB>(), 0, op.stream>>>(a);
    MFG_HIP(hipGetLastError());
  });
}

// R = residual(u, acc) (+ the interface traction over n_faces stored face
// pairs) with constrained entries zero; returns ||R||.  One host round trip:
// the norm, the det F flag and (sc) the reducer's scalar block come back
// together.
double solid_residual(const OperatorData &op, const double *u, const double *acc, const double *body, double *R,
                      Reducer &red, int64_t n_faces = 0, const int32_t *faces = nullptr,
                      const double *traction = nullptr, double *sc = nullptr) {
  MFG_REQUIRE(op.kind == 0 && op.precision == MFG_FP64, MFG_ERR_ARGUMENT, "residual needs an FP64 solid operator");
  vec_fill(MFG_FP64, R, op.n_dofs, 0.0, op.stream);
  vec_fill(MFG_FP32, op.flag.ptr, 1, 0.0, op.stream);  // int 0 == float 0 bit pattern
  switch (op.p) {
    case 1: residual_launch<1>(op, u, acc, body, R); break;
    case 2: residual_launch<2>(op, u, acc, body, R); break;
    case 3: residual_launch<3>(op, u, acc, body, R); break;
    case 4: residual_launch<4>(op, u, acc, body, R); break;
    case 5: residual_launch<5>(op, u, acc, body, R); break;
    case 6: residual_launch<6>(op, u, acc, body, R); break;
    default: throw Error(MFG_ERR_ARGUMENT, "unsupported degree");
  }
  // mf_elasticity.h:1849-1897, skipping constrained dofs (zero after :1907)
  add_interface_traction(op, n_faces, faces, u, traction, R);
  red.dot_dev(MFG_FP64, R, R, op.n_dofs, kRedSet, kSlotNewton, op.stream);
  int h = 0;
  MFG_HIP(hipMemcpyAsync(&h, op.flag.ptr, sizeof(int), hipMemcpyDeviceToHost, op.stream));
  double local[kRedScalars];
  if (!sc) sc = local;
  red.read(sc, op.stream);
  MFG_REQUIRE(h == 0, MFG_ERR_NUMERIC, "det_F is not positive");
  return std::sqrt(sc[kSlotNewton]);
}

void SolidProblem::set_stream(hipStream_t s) {
  stream = s;
  for (auto &o : levels) o->stream = s;
  if (fine64) fine64->stream = s;
}

// update_matrix_free (mf_elasticity.h:1289-1353): caches and diagonals on
// every level at the interpolated displacement, then the GMG set-up.
void SolidProblem::update_matrix_free() {
  OperatorData &F = fine();
  F.cache(u.ptr, MFG_FP64);
  if (nm.linear_preconditioner != MFG_NEWTON_GMG) {
    // setup_operator_cache of the fine operator only (mf_elasticity.h:1320-1324)
    if (nm.linear_preconditioner == MFG_NEWTON_JACOBI) F.compute_diagonal();
    return;
  }
  const int top = L - 1;
  if (fine64) {
    vec_convert(levels[top]->precision, u_lev[top].ptr, MFG_FP64, u.ptr, n(), stream);
    levels[top]->cache(u_lev[top].ptr, levels[top]->precision);
  }
  const void *u_top = fine64 ? u_lev[top].ptr : u.ptr;
  for (int l = top; l >= 1; --l) {
    const void *uf = l == top ? u_top : u_lev[l].ptr;
    transfers[l - 1]->interpolate(u_lev[l - 1].ptr, uf);
    levels[l - 1]->cache(u_lev[l - 1].ptr, levels[l - 1]->precision);
  }
  mg->setup();
}

void SolidProblem::step(mfg_step_info &info) {
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t N = n();
  OperatorData &F = fine();
  std::memset(&info, 0, sizeof(info));
  const int grid = (int)std::min<int64_t>((N + 255) / 256, 8192);
  vec_fill(MFG_FP64, delta_u.ptr, N, 0.0, stream);
  double res0 = 1.0, upd = 1.0, upd0 = 1.0, res = 0.0;
  bool converged = false;
  info.condition_estimate = 1.0;
  double sc[kRedScalars];
  int it = 0;
  for (; it < nm.max_newton; ++it) {
    newmark_acc_kernel<double><<<grid, 256, 0, stream>>>(a.as<double>(), delta_u.as<double>(), v_old.as<double>(),
                                                         a_old.as<double>(), a1, a2, a3, N);
    MFG_HIP(hipGetLastError());
    res = solid_residual(F, u.as<double>(), a.as<double>(), nm.body_force, R.as<double>(), red, iface_n,
                         iface_faces.as<int32_t>(), iface_traction, sc);
    if (it > 0) {  // ||du|| of the previous update came back with the residual
      upd = std::sqrt(sc[kSlotUpd]);
      if (it == 1) upd0 = upd;
    }
    if (it == 0) {
      res0 = res;
      info.residual_0 = res;
    }
    // check_convergence (mf_elasticity.h:1357-1395)
    if (it >= nm.min_newton) {
      const double rn = res0 != 0.0 ? res / res0 : res, un = upd0 != 0.0 ? upd / upd0 : upd;
      if (res <= nm.tol_f_abs || upd <= nm.tol_u_abs || (rn <= nm.tol_f && un <= nm.tol_u)) {
        converged = true;
        break;
      }
    }
    update_matrix_free();
    ws.ensure((size_t)N * 8);
    if (nm.cond_cg_iterations > 0) {
      // solve_linear_system's condition estimate (mf_elasticity.h:1998-2013):
      // 20 unpreconditioned CG steps that only report the Lanczos condition
      // number.  With FP32 levels it runs on the FP32 fine-level operator, like
      // the rest of the mixed-precision preconditioning work; the solution
      // does not depend on it.
      const CGVectors cv{ws.r.ptr, ws.z.ptr, ws.p.ptr, ws.h.ptr};
      if (fine64 && nm.linear_preconditioner == MFG_NEWTON_GMG) {
        OperatorData &F32 = *levels.back();
        if (cond_b.bytes != (size_t)N * 4) {
          cond_b.allocate((size_t)N * 4);
          cond_x.allocate((size_t)N * 4);
        }
        vec_convert(MFG_FP32, cond_b.ptr, MFG_FP64, R.ptr, N, stream);
        info.condition_estimate =
            cg_condition_estimate(F32, cond_b.ptr, cond_x.ptr, nm.cond_cg_iterations, nm.tol_lin * res, cv, red);
      } else {
        info.condition_estimate =
            cg_condition_estimate(F, R.ptr, du.ptr, nm.cond_cg_iterations, nm.tol_lin * res, cv, red);
      }
    }
    vec_fill(MFG_FP64, du.ptr, N, 0.0, stream);
    mfg_solve_info si{};
    if (nm.linear_preconditioner == MFG_NEWTON_JACOBI)
      cg_solve(F, MFG_PRECOND_JACOBI, nullptr, nm.jacobi_relaxation, R.ptr, du.ptr, nm.tol_lin, nm.max_lin, &si, ws,
               red);
    else if (nm.linear_preconditioner == MFG_NEWTON_NONE)
      cg_solve(F, MFG_PRECOND_NONE, nullptr, 1.0, R.ptr, du.ptr, nm.tol_lin, nm.max_lin, &si, ws, red);
    else
      cg_solve(F, MFG_PRECOND_MG, mg.get(), 1.0, R.ptr, du.ptr, nm.tol_lin, nm.max_lin, &si, ws, red);
    info.linear_iterations += si.iterations;
    red.dot_dev(MFG_FP64, du.ptr, du.ptr, N, kRedSet, kSlotUpd, stream);
    add2_kernel<double><<<grid, 256, 0, stream>>>(delta_u.as<double>(), u.as<double>(), du.as<double>(), N);
    MFG_HIP(hipGetLastError());
  }
  if (it == nm.max_newton && it > 0) {  // the last update's norm was not read with a residual
    red.read(sc, stream);
    upd = std::sqrt(sc[kSlotUpd]);
    if (it == 1) upd0 = upd;
  }
  info.newton_iterations = it;
  info.residual_final = res;
  info.update_final = upd;
  info.converged = converged;
  if (converged) {
    // update_velocity (:1600-1609): v = a4 du + a5 v_old + a6 a_old; then
    // v_old <- v, a_old <- a
    newmark_vel_kernel<double><<<grid, 256, 0, stream>>>(v.as<double>(), delta_u.as<double>(), v_old.as<double>(),
                                                         a_old.as<double>(), a4, a5, a6, N);
    MFG_HIP(hipGetLastError());
    vec_convert(MFG_FP64, v_old.ptr, MFG_FP64, v.ptr, N, stream);
    vec_convert(MFG_FP64, a_old.ptr, MFG_FP64, a.ptr, N, stream);
  }
  MFG_HIP(hipStreamSynchronize(stream));
  info.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (!converged) throw Error(MFG_ERR_NOCONV, "No convergence in nonlinear solver!");
}

size_t SolidProblem::memory() const {
  size_t b = u.bytes + du.bytes + delta_u.bytes + v.bytes + a.bytes + v_old.bytes + a_old.bytes + R.bytes +
             iface_faces.bytes;
  for (auto &o : levels) b += o->memory_consumption() + o->diag.bytes + o->inv_diag.bytes + o->vertices.bytes;
  for (auto &L : mg->levels)
    for (const DeviceBuffer *x : {&L.sol, &L.def, &L.tmp, &L.d, &L.t, &L.cg_r, &L.cg_p, &L.cg_h, &L.cg_z}) b += x->bytes;
  if (fine64) b += fine64->memory_consumption() + fine64->vertices.bytes;
  for (auto &t : transfers) b += t->lattice.bytes;
  for (auto &x : u_lev) b += x.bytes;
  b += ws.r.bytes + ws.z.bytes + ws.p.bytes + ws.h.bytes + cond_b.bytes + cond_x.bytes;
  return b;
}

size_t estimate_solid_problem_bytes(int n_levels, const mfg_level_desc *levels, int level_precision) {
  size_t b = 0;
  for (int l = 0; l < n_levels; ++l) {
    const mfg_level_desc &d = levels[l];
    const int64_t n3 = (int64_t)(d.degree + 1) * (d.degree + 1) * (d.degree + 1);
    const int64_t nqp = (int64_t)d.n_cells * n3, ndof = d.n_nodes * d.n_components;
    const size_t es = (l == 0 || level_precision == MFG_FP64) ? 8 : 4;
    b += nqp * (18 * es + 4) + (size_t)d.n_cells * 24 * 8 + 11 * ndof * es;  // caches, map, vertices, diag + MG vectors
    if (nqp <= kGatherMaxPoints) b += gather_bytes(nqp, d.n_nodes, d.n_components, es);
    if (l > 0) b += (size_t)levels[l - 1].n_cells * (2 * d.degree + 1) * (2 * d.degree + 1) * (2 * d.degree + 1) * 4;
    if (l == n_levels - 1) {
      b += 12 * ndof * 8;  // state, residual, CG workspace
      if (level_precision == MFG_FP32 && n_levels > 1) {
        b += nqp * (18 * 8 + 4) + (size_t)d.n_cells * 24 * 8 + 2 * ndof * 4;  // + FP32 condition-estimate vectors
        if (nqp <= kGatherMaxPoints) b += gather_bytes(nqp, d.n_nodes, d.n_components, 8);
      }
    }
  }
  return b;
}

std::unique_ptr<SolidProblem> make_solid_problem(int n_levels, const mfg_level_desc *levels,
                                                 const int32_t *const *children, const mfg_solid_params &material,
                                                 const mfg_newmark_params &newmark, const mfg_mg_params &mgp,
                                                 int level_precision, hipStream_t stream) {
  MFG_REQUIRE(n_levels >= 1 && levels, MFG_ERR_ARGUMENT, "null argument");
  MFG_REQUIRE(n_levels == 1 || children, MFG_ERR_ARGUMENT, "children tables required");
  MFG_REQUIRE(level_precision == MFG_FP64 || level_precision == MFG_FP32, MFG_ERR_ARGUMENT, "bad precision");
  MFG_REQUIRE(newmark.dt > 0 && newmark.beta > 0, MFG_ERR_ARGUMENT, "bad Newmark parameters");
  MFG_REQUIRE(newmark.linear_preconditioner >= MFG_NEWTON_GMG && newmark.linear_preconditioner <= MFG_NEWTON_NONE,
              MFG_ERR_ARGUMENT, "Preconditioner type not implemented");
  MFG_REQUIRE(newmark.linear_preconditioner != MFG_NEWTON_JACOBI || newmark.jacobi_relaxation > 0, MFG_ERR_ARGUMENT,
              "the Jacobi preconditioner needs a positive relaxation");
  MFG_REQUIRE(material.mu > 0 && material.nu > -1 && material.nu < 0.5 && material.rho >= 0, MFG_ERR_ARGUMENT,
              "bad material");
  MFG_REQUIRE(material.inclusion_mu == 0 || (material.inclusion_mu > 0 && material.inclusion_nu > -1 &&
                                             material.inclusion_nu < 0.5 && material.inclusion_rho >= 0),
              MFG_ERR_ARGUMENT, "bad inclusion material");
  MFG_REQUIRE(material.formulation == 0 || material.formulation == 1, MFG_ERR_ARGUMENT,
              "Unknown material formulation");
  MFG_REQUIRE(material.caching >= MFG_CACHING_SCALAR_REFERENTIAL && material.caching <= MFG_CACHING_TENSOR4_NS,
              MFG_ERR_ARGUMENT, "Unknown caching strategy");
  MFG_REQUIRE(material.formulation == 1 || material.caching == MFG_CACHING_TENSOR2, MFG_ERR_ARGUMENT,
              "formulation 0 needs the tensor2 caching");
  auto p = std::make_unique<SolidProblem>();
  p->L = n_levels;
  p->level_precision = level_precision;
  p->nm = newmark;
  const double beta = newmark.beta, gamma = newmark.gamma, dt = newmark.dt;
  // mf_elasticity.h:240-245
  p->a1 = 1.0 / (beta * dt * dt);
  p->a2 = 1.0 / (beta * dt);
  p->a3 = (1 - 2 * beta) / (2 * beta);
  p->a4 = gamma / (beta * dt);
  p->a5 = 1 - gamma / beta;
  p->a6 = (1 - gamma / (2 * beta)) * dt;
  mfg_solid_params sp = material;
  sp.alpha_1 = p->a1;

  auto make_op = [&](const mfg_level_desc &d, int prec) {
    auto op = std::make_unique<mfg_operator_t>();
    op->kind = 0;
    op->precision = prec;
    op->sp = sp;
    op->stream = stream;
    op->setup_layout(d);
    // the problem's operators never enter a batched launch, so its levels on
    // affine cells take the 10-word point storage (faster and smaller than the
    // folded 16 words, in FP64 and FP32)
    if (sp.formulation == 1 && sp.caching == MFG_CACHING_TENSOR2 && op->affine_ref) op->t2_storage = MFG_T2_GRADIENT;
    op->setup_solid();
    return op;
  };
  for (int l = 0; l < n_levels; ++l) p->levels.push_back(make_op(levels[l], l == 0 ? MFG_FP64 : level_precision));
  if (level_precision == MFG_FP32 && n_levels > 1) p->fine64 = make_op(levels[n_levels - 1], MFG_FP64);
  for (int l = 1; l < n_levels; ++l) {
    auto t = std::make_unique<mfg_transfer_t>();
    t->build(p->levels[l - 1].get(), p->levels[l].get(), children[l - 1]);
    p->transfers.push_back(std::move(t));
  }
  p->mg = std::make_unique<mfg_mg_t>();
  p->mg->prm = mgp;
  for (int l = 0; l < n_levels; ++l) {
    MGLevel Lv;
    Lv.op = p->levels[l].get();
    p->mg->levels.push_back(std::move(Lv));
    if (l > 0) p->mg->transfers.push_back(p->transfers[l - 1].get());
  }
  p->stream = stream;
  const int64_t N = p->fine().n_dofs;
  for (DeviceBuffer *b : {&p->u, &p->du, &p->delta_u, &p->v, &p->a, &p->v_old, &p->a_old, &p->R}) {
    b->allocate(N * 8);
    vec_fill(MFG_FP64, b->ptr, N, 0.0, stream);
  }
  p->u_lev.resize(n_levels);
  for (int l = 0; l < n_levels; ++l) {
    if (l == n_levels - 1 && !p->fine64) continue;  // the FP64 active level uses u itself
    p->u_lev[l].allocate(p->levels[l]->n_dofs * p->levels[l]->elem_size());
  }
  MFG_HIP(hipStreamSynchronize(stream));
  return p;
}

}  // namespace mfg

using namespace mfg;



extern "C" {

int mfg_solid_residual(mfg_operator op, const double *u, const double *acc, const double body[3], double *R,
                       double *norm) {
  return guarded([&] {
    MFG_REQUIRE(op && u && acc && body && R, MFG_ERR_ARGUMENT, "null argument");
    static thread_local Reducer red;
    const double n = solid_residual(*op, u, acc, body, R, red);
    if (norm) *norm = n;
  });
}

int mfg_solid_problem_create(int32_t n_levels, const mfg_level_desc *levels, const int32_t *const *children,
                             const mfg_solid_params *material, const mfg_newmark_params *newmark,
                             const mfg_mg_params *mgp, int32_t level_precision, mfg_solid_problem *out) {
  return guarded([&] {
    MFG_REQUIRE(levels && material && newmark && mgp && out, MFG_ERR_ARGUMENT, "null argument");
    auto h = std::make_unique<mfg_solid_problem_t>();
    h->p = make_solid_problem(n_levels, levels, children, *material, *newmark, *mgp, level_precision, nullptr);
    *out = h.release();
  });
}

int mfg_solid_problem_destroy(mfg_solid_problem p) {
  return guarded([&] { delete p; });
}

int mfg_solid_problem_set_stream(mfg_solid_problem p, void *s) {
  return guarded([&] {
    MFG_REQUIRE(p, MFG_ERR_ARGUMENT, "null problem");
    p->p->set_stream((hipStream_t)s);
  });
}

int mfg_solid_problem_set_state(mfg_solid_problem p, const double *u, const double *v, const double *a) {
  return guarded([&] {
    MFG_REQUIRE(p, MFG_ERR_ARGUMENT, "null problem");
    SolidProblem &q = *p->p;
    const int64_t N = q.n();
    if (u) vec_convert(MFG_FP64, q.u.ptr, MFG_FP64, u, N, q.stream);
    if (v) vec_convert(MFG_FP64, q.v_old.ptr, MFG_FP64, v, N, q.stream);
    if (a) vec_convert(MFG_FP64, q.a_old.ptr, MFG_FP64, a, N, q.stream);
    if (v) vec_convert(MFG_FP64, q.v.ptr, MFG_FP64, v, N, q.stream);
    if (a) vec_convert(MFG_FP64, q.a.ptr, MFG_FP64, a, N, q.stream);
  });
}

int mfg_solid_problem_get_state(mfg_solid_problem p, double *u, double *v, double *a) {
  return guarded([&] {
    MFG_REQUIRE(p, MFG_ERR_ARGUMENT, "null problem");
    SolidProblem &q = *p->p;
    const int64_t N = q.n();
    if (u) vec_convert(MFG_FP64, u, MFG_FP64, q.u.ptr, N, q.stream);
    if (v) vec_convert(MFG_FP64, v, MFG_FP64, q.v_old.ptr, N, q.stream);
    if (a) vec_convert(MFG_FP64, a, MFG_FP64, q.a_old.ptr, N, q.stream);
  });
}

int mfg_solid_problem_step(mfg_solid_problem p, mfg_step_info *info) {
  mfg_step_info dummy;
  if (!info) info = &dummy;
  return guarded([&] {
    MFG_REQUIRE(p, MFG_ERR_ARGUMENT, "null problem");
    p->p->step(*info);
  });
}

int mfg_solid_problem_set_interface(mfg_solid_problem p, int64_t n_faces, const int32_t *faces,
                                    const double *traction) {
  return guarded([&] {
    MFG_REQUIRE(p, MFG_ERR_ARGUMENT, "null problem");
    MFG_REQUIRE(n_faces == 0 || traction, MFG_ERR_ARGUMENT, "null traction");
    SolidProblem &sp = *p->p;
    DeviceBuffer fb = stored_faces(sp.fine(), n_faces, faces);
    sp.iface_faces = std::move(fb);
    sp.iface_n = n_faces;
    sp.iface_traction = n_faces ? traction : nullptr;
  });
}

size_t mfg_solid_problem_memory(mfg_solid_problem p) { return p ? p->p->memory() : 0; }

}  // extern "C"
