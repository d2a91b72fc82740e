// Host-side objects of the multigrid preconditioner and CG solver.
#pragma once
#include "common.hpp"
#include "operator.hpp"

namespace mfg {

// Deterministic device reductions (fixed grid, fixed-order sum).  The CG
// forms keep their results and the step lengths derived from them in a block
// of device scalars (RedSlot); the host reads that block back when it has a
// decision to make.
struct Reducer {
  DeviceBuffer own;            // partials and scalars, when the reducer allocates them
  DeviceBuffer hist_buf;       // step lengths of fixed-count runs (alphas, then betas)
  double *partials = nullptr;  // kRedBlocks
  double *scal = nullptr;      // kRedScalars
  int hist_cap = 0;
  Reducer();
  // Memory provided by the caller (kRedBlocks + kRedScalars doubles): the
  // reducer allocates nothing.
  Reducer(double *partials, double *scalars);
  Reducer(const Reducer &) = delete;
  Reducer &operator=(const Reducer &) = delete;
  // a.b, read back at once
  double dot(int prec, const void *a, const void *b, int64_t n, hipStream_t s);
  // a.b into the scalar block through the epilogue `mode` (RedMode)
  void dot_dev(int prec, const void *a, const void *b, int64_t n, int mode, int slot, hipStream_t s,
               double tol = 0, double rel = 0, bool record = false);
  // x += alpha p; r -= alpha h (alpha from the scalar block); r.r through `mode`
  void cg_update_dev(int prec, void *x, void *r, const void *p, const void *h, int64_t n, int mode, hipStream_t s);
  // z = omega D^{-1} r, and r.z through `mode` (one pass over the vectors)
  void jacobi_dot_dev(int prec, void *z, const void *invD, const void *r, double omega, int64_t n, int mode,
                      hipStream_t s, bool record = false);
  // the partial sums of the last reduction through another epilogue (a second
  // dot product of the same vectors, without reading them again)
  void repeat_epilogue(int mode, int slot, hipStream_t s, double tol = 0, double rel = 0, bool record = false);
  // p = z + beta p (beta from the scalar block)
  void xpby_dev(int prec, void *p, const void *z, int64_t n, hipStream_t s);
  // the scalar block, synchronising the stream
  void read(double *out, hipStream_t s);
  void ensure_hist(int cap);
  // recorded step lengths of the last fixed-count run (the counts from `sc`, a block read)
  void read_hist(const double *sc, std::vector<double> &alphas, std::vector<double> &betas, hipStream_t s);

 private:
  void epilogue(int mode, int slot, double tol, double rel, bool record, hipStream_t s);
};

void vec_fill(int prec, void *x, int64_t n, double v, hipStream_t s);
void vec_convert(int prec_dst, void *dst, int prec_src, const void *src, int64_t n, hipStream_t s);
void vec_xpby(int prec, void *p, const void *d, double beta, int64_t n, hipStream_t s);
void vec_residual(int prec, void *t, const void *b, int64_t n, hipStream_t s);
void vec_add_scaled(int prec, void *y, const void *a, double sc, const void *b, int64_t n, hipStream_t s);

struct TransferData {
  OperatorData *coarse = nullptr, *fine = nullptr;
  int p = 1, nc = 1;
  DeviceBuffer lattice;  // (2p+1)^3 entries per coarse cell (stored order)
  DeviceBuffer E, I;     // embedding (2p+1)x(p+1), interpolation (p+1)x(2p+1)
  void build(OperatorData *c, OperatorData *f, const int32_t *children);
  void prolongate_add(void *fine, const void *coarse);
  void restrict_add(void *coarse, const void *fine);
  void interpolate(void *coarse, const void *fine);

 private:
  void run(int kind, void *fine, const void *coarse_in, void *coarse_out, const void *fine_in);
};

struct MGLevel {
  OperatorData *op = nullptr;
  DeviceBuffer sol, def, tmp, d, t, cg_r, cg_p, cg_h, cg_z;
  int degree = 5;
  double theta = 1, delta = 0, lmin = 0, lmax = 0;
  void allocate();
  void estimate_eigenvalues(int n_it, Reducer &red);
  void setup_chebyshev(int deg, double range, int eig_it, Reducer &red);
  void chebyshev(void *x, const void *b, bool zero_guess);
};

struct MGData {
  std::vector<MGLevel> levels;
  std::vector<TransferData *> transfers;
  mfg_mg_params prm{};
  Reducer red;
  bool ready = false;
  // A single-level hierarchy is its own coarse solve; an inner CG would make
  // the preconditioner non-linear (outer CG stagnates), so the Chebyshev
  // solver mode is used there whatever prm.coarse_kind says.
  int coarse_kind() const { return levels.size() == 1 ? 0 : prm.coarse_kind; }
  // active iterations of the last coarse CG: the length of the next one's
  // first run between read-backs
  int coarse_iterations = 0;
  void setup();
  void vmult(void *dst, const void *src, int vec_prec);
  void v_cycle();
  void coarse_solve();
  // Multigrid::level_v_step unrolled: pre-smoothing, residual and restriction
  // from the finest level down to level 1, then (after the coarse solve)
  // prolongation and post-smoothing back up.
  void down_sweep();
  void up_sweep();
  // Optionally the two sweeps are captured into hipGraphs on the second
  // V-cycle and replayed on a stream of their own; the coarse solve, which may
  // synchronise with the host, runs eagerly between the two replays.  A graph
  // is re-captured after setup() or when a level operator reallocated a
  // buffer it reads.  Off by default: the small levels are bound by each
  // colour launch's own critical path on the device, not by launch cost, and
  // the re-capture after every setup() adds latency.
  bool use_graph = false;
  hipStream_t graph_stream = nullptr;
  hipEvent_t graph_in = nullptr, graph_out = nullptr;
  hipGraphExec_t graph_down = nullptr, graph_up = nullptr;
  std::vector<uint64_t> graph_versions;
  int graph_calls = 0;
  size_t graph_nodes = 0;
  bool graphs_allowed() const;
  void run_graphed(hipStream_t caller);
  void drop_graphs();
  MGData() = default;
  MGData(const MGData &) = delete;
  MGData &operator=(const MGData &) = delete;
  ~MGData();
};

struct CGWorkspace {
  DeviceBuffer r, z, p, h;
  void ensure(size_t bytes);
};

// The four CG vectors, wherever they live.
struct CGVectors {
  void *r, *z, *p, *h;
};

// Extreme eigenvalues of the Lanczos matrix of CG coefficients.
void lanczos_estimate(const std::vector<double> &alphas, const std::vector<double> &betas, double &lmin,
                      double &lmax);

void cg_solve(const OperatorData &A, int precond, MGData *mg, double omega, const void *b, void *x, double tol_rel,
              int max_iter, mfg_solve_info *info, const CGVectors &v, Reducer &red);
inline void cg_solve(const OperatorData &A, int precond, MGData *mg, double omega, const void *b, void *x,
                     double tol_rel, int max_iter, mfg_solve_info *info, CGWorkspace &ws, Reducer &red) {
  ws.ensure((size_t)A.n_dofs * A.elem_size());
  cg_solve(A, precond, mg, omega, b, x, tol_rel, max_iter, info, CGVectors{ws.r.ptr, ws.z.ptr, ws.p.ptr, ws.h.ptr},
           red);
}

// Unpreconditioned CG from x = 0 for at most n_it iterations or until
// ||r|| <= tol (deal.II's IterationNumberControl), run without host round
// trips; returns the condition estimate lmax / lmin of its Lanczos matrix
// (1 when no step was taken).  x is left at the iterate; the vectors are
// scratch.
double cg_condition_estimate(const OperatorData &A, const void *b, void *x, int n_it, double tol, const CGVectors &v,
                             Reducer &red);

}  // namespace mfg

struct mfg_transfer_t : mfg::TransferData {};
struct mfg_mg_t : mfg::MGData {};
