/*
 * CPU oracle: a host restatement of the reference's matrix-free operator and
 * MG-preconditioned CG path, written from the mathematics the reference code
 * states (mf_nh_operator.h, material.h, heat_transfer.h, mf_elasticity.h).
 *
 * This library is TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker (or the
 * timed CPU baseline); the product path never calls into it.
 *
 * Pinning status of every function is stated next to its declaration:
 *   [pinned: ...] names the independent check(s) in tests/ that would fail on
 *   a dropped term, a wrong sign/index or a transposed operand.
 *
 * All arrays are host arrays.  DoF numbering: dof = n_components*node + comp,
 * nodes numbered lexicographically over the structured box
 * (I + NX*(J + NY*K)), cells lexicographic (cx + nx*(cy + ny*cz)), cell-local
 * node order lexicographic (a + N*(b + N*c)), N = degree+1.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- 1D FE --- */
/* Gauss-Lobatto support points of FE_Q(p) on [0,1] (p+1 points).
 * [pinned: closed forms for p<=4, endpoint/symmetry, Lagrange partition of
 *  unity; tests/test_oracle_fe.py] */
void orc_gll_points(int p, double *x);
/* QGauss(n) points/weights on [0,1].
 * [pinned: closed forms n<=3, exact integration of x^k, k<2n] */
void orc_gauss_rule(int n, double *x, double *w);
/* Lagrange basis on nodes z[0..m-1] evaluated at x[0..nx-1]:
 * val[i*m+a] = l_a(x_i), der[i*m+a] = l_a'(x_i).
 * [pinned: derivative vs. finite differences, reproduction of monomials] */
void orc_lagrange(int m, const double *z, int nx, const double *x, double *val,
                  double *der);

/* --------------------------------------------------------------- meshes --- */
/* Structured box [lo,hi] with n[0]xn[1]xn[2] hex cells, FE_Q(p)^nc on GLL
 * nodes.  Dirichlet faces: bit f of face_mask selects face f in the
 * deal.II colorize numbering (0:x=lo,1:x=hi,2:y=lo,3:y=hi,4:z=lo,5:z=hi);
 * comp_mask selects constrained components.
 * [pinned: closed-form node coordinates and constrained counts
 *  3*(ny*p+1)*(nz*p+1); tests/test_abi.py, tests/test_golden.py] */
void *orc_mesh_create(const double *lo, const double *hi, const int *n, int p,
                      int nc, int face_mask, int comp_mask);
/* The same with a component mask per face (face_comp[6]); a node on several
 * faces is constrained in the union of their masks.
 * [pinned: per-face constrained counts in closed form, bit-identical to the
 *  device builder; tests/test_turek.py] */
void *orc_mesh_create_bc(const double *lo, const double *hi, const int *n, int p,
                         int nc, const int *face_comp);
void orc_mesh_destroy(void *m);
/* out[0..5] = n_cells, n_nodes, n_dofs, n_constrained, nodes_per_cell, nc */
void orc_mesh_sizes(void *m, int64_t *out);
void orc_mesh_cell_nodes(void *m, int32_t *out);    /* n_cells*(p+1)^3 */
void orc_mesh_constrained(void *m, int32_t *out);   /* sorted dof list */
void orc_mesh_node_coords(void *m, double *out);    /* n_nodes*3 */
void orc_mesh_cell_vertices(void *m, double *out);  /* n_cells*8*3 */
/* new cell vertices; nodes move to the trilinear images of their GLL points
 * (call before creating operators on the mesh) */
void orc_mesh_set_cell_vertices(void *m, const double *v);

/* Nodal interpolation of the function u0(x)=amp*sin(pi x)sin(pi y)sin(pi z)
 * per component plus noise*U[-1,1] from splitmix64(seed, dof); constrained
 * entries zeroed.  Used to generate the linearisation point. */
void orc_fill_u0(void *m, double amp, double noise, uint64_t seed, double *u);
/* splitmix64-based U[-1,1] vector of length n (constrained-agnostic). */
void orc_fill_uniform(uint64_t seed, int64_t n, double *v);

/* ------------------------------------------------------ solid operator --- */
/* caching: 0 scalar_referential, 1 tensor2, 2 tensor4, 3 tensor4_ns
 * formulation: 0 (volumetric/isochoric split) or 1 (one-field, default).
 * Material: mu, nu, rho (material.h:78-94); alpha_1 = 1/(beta dt^2).
 * [pinned: brute-force element assembly (orc_solid_vmult_bruteforce),
 *  agreement of the four cachings, u=0 reduction to linear elasticity with
 *  Lame pair (mu, 2 lambda) + rho*alpha_1*M (closed-form mass 1^T M 1),
 *  symmetry, finite differences of orc_solid_residual] */
void *orc_solid_create(void *mesh, int n_q, double mu, double nu, double rho,
                       double alpha_1, int formulation, int caching);
void orc_solid_destroy(void *op);
/* set_material's inclusion material (mf_nh_operator.h:121-128, 529-539):
 * cells whose material_ids entry (one per cell, NULL: none) is not 0 use
 * (mu, nu, rho), the others the material of orc_solid_create
 * (mf_nh_operator.h:404, 658; residual mf_elasticity.h:1781-1784).
 * [pinned: swapping the two materials cell by cell and adding gives the sum
 * of the two one-material operators (and residuals); all cells of id 1 give
 * the one-material operator of the inclusion; tests/test_oracle_operators.py] */
void orc_solid_set_inclusion(void *op, const int32_t *material_ids, double mu, double nu, double rho);
/* cache(): linearise at displacement u (mf_nh_operator.h:392-523). Returns 0,
 * or -1 if det F <= 0 at some quadrature point. */
int orc_solid_cache(void *op, const double *u);
/* dst = A src with identity rows on constrained dofs (mf_nh_operator.h:575-625) */
void orc_solid_vmult(void *op, double *dst, const double *src);
/* Independent dense element-matrix path (no sum factorisation, no cache):
 * F, tau and the tangent are recomputed from u with plain quadrature loops. */
void orc_solid_vmult_bruteforce(void *op, double *dst, const double *src);
/* diagonal with constrained entries = 1 (mf_nh_operator.h:1023-1060) */
void orc_solid_diagonal(void *op, double *diag);
/* tensor2 formulation-1 quadrature data in the device layout:
 * out[f*n_qp + cell*nq3 + q], f = 0..17: Jinv_cur[f0][e] row-major (9),
 * JxW_cur, c2 = 2(mu-2 lambda lnJ)/J, c1 = 1/J, T=tau/J (xx,yy,zz,xy,xz,yz) */
void orc_solid_tensor2_data(void *op, double *out);
/* Residual R(u, a) = -int tau F^{-T}:Grad w + int rho (b - a).w, constrained
 * entries zeroed (mf_elasticity.h:1764-1846, 1905-1907).  body = constant
 * body force (3). */
void orc_solid_residual(void *op, const double *u, const double *acc,
                        const double *body, double *R);

/* Interface traction (mf_elasticity.h:1849-1897): rhs += int t0 . w dA over
 * n_faces (cell, face = 2 axis + side) pairs, t0 = |det F F^{-T} N| t at the
 * displacement u, computed through Nanson's formula from the deformed
 * tangents; t: 3 per face point, point qa + N qb over the Gauss points of the
 * two tangential directions (ascending); constrained dofs untouched. */
void orc_interface_traction(void *mesh, int64_t n_faces, const int32_t *faces, const double *u, const double *t,
                            double *rhs);

/* Field values (mesh components per node) at the face points, point order as
 * orc_interface_traction; out: n_faces * N^2 * nc. */
void orc_face_values(void *mesh, int64_t n_faces, const int32_t *faces, const double *u, double *out);
/* Heat interface flux (heat_transfer.h:618-652): rhs -= dt int q phi dA, q per
 * face point; constrained dofs untouched (scalar meshes). */
void orc_interface_flux(void *mesh, int64_t n_faces, const int32_t *faces, const double *flux, double dt,
                        double *rhs);

/* ------------------------------------------------------- heat operator --- */
/* A = M + dt * c(x) K on scalar FE_Q(p) (heat_transfer.h:234-250).
 * coef: per (cell, q) conductivity table, or NULL for c == 1.
 * [pinned: brute force element assembly; closed form 1^T M 1 = |Omega|;
 *  exact reproduction of a quadratic manufactured solution] */
void *orc_heat_create(void *mesh, int n_q, double dt, const double *coef);
void orc_heat_destroy(void *op);
void orc_heat_vmult(void *op, double *dst, const double *src);
void orc_heat_vmult_bruteforce(void *op, double *dst, const double *src);
void orc_heat_diagonal(void *op, double *diag);
/* Unconstrained apply (no constraint handling at all): used for the
 * inhomogeneous-Dirichlet lifting of the right-hand side. */
void orc_heat_vmult_noconstraints(void *op, double *dst, const double *src);
/* mesh an operator was created on (kind 0 solid, 1 heat) */
void *orc_operator_mesh(int kind, void *op);
/* load vector: int f phi_i dx for constant f */
void orc_heat_load(void *op, double f, double *b);

/* ------------------------------------------------------- MG hierarchy --- */
/* Transfer between nested boxes (fine = coarse refined once).
 * prolongate_add: fine += P coarse (coarse constrained read as 0, fine
 * constrained untouched); restrict_add: coarse += P^T fine (fine constrained
 * read as 0, coarse constrained untouched); interpolate: coarse = nodal
 * interpolant of the fine FE function.
 * [pinned: polynomial reproduction, <P c, f> = <c, P^T f>, interpolate o
 *  prolongate = identity on coarse space] */
void orc_prolongate_add(void *coarse_mesh, void *fine_mesh, const double *c,
                        double *f);
void orc_restrict_add(void *coarse_mesh, void *fine_mesh, const double *f,
                      double *c);
void orc_interpolate(void *coarse_mesh, void *fine_mesh, const double *f,
                     double *c);

/* MG-preconditioned CG on a level hierarchy of solid or heat operators.
 * ops[l], meshes[l] for l = 0..L (coarse..fine).  kind: 0 solid, 1 heat.
 * Chebyshev: degree, smoothing_range, eig_cg iterations on levels >0;
 * coarse: 0 -> Chebyshev "as solver" (range 1e-3, degree from the estimate,
 * eig iterations = m()), 1 -> CG to relative 1e-3 preconditioned by the
 * level-0 Chebyshev.  Returns iterations (or -1 on no convergence).
 * x is the initial guess on entry (zeroed by the caller in practice).
 * [pinned: convergence to the requested relative residual, rhs = A x_true
 *  recovers x_true; Lanczos eigenvalue bounds vs. dense eigen-solve on tiny
 *  levels; orc_mg_vmult with the Chebyshev coarse solver equals a V-cycle
 *  composed independently in numpy from dense level matrices, dense transfer
 *  and the closed-form Chebyshev polynomials, and is a symmetric operator;
 *  tests/test_oracle_mg.py] */
typedef struct {
  int cheb_degree;
  double smoothing_range;
  int eig_cg_iterations;
  int coarse_kind;
  double tol_rel;
  int max_iter;
} orc_mg_params;
int orc_mg_cg_solve(int kind, int n_levels, void **meshes, void **ops,
                    const orc_mg_params *prm, const double *b, double *x,
                    double *final_rel_residual);
/* One preconditioner application dst = V-cycle(src) after the same set-up;
 * eig (optional): per level lmax, lmin, Chebyshev degree. */
void orc_mg_vmult(int kind, int n_levels, void **meshes, void **ops,
                  const orc_mg_params *prm, const double *src, double *dst,
                  double *eig);
/* One Newmark step, Newton + MG-CG (solid, formulation of the level ops).
 * nm = {beta, gamma, dt, bx, by, bz, tol_f, tol_u, tol_f_abs, tol_u_abs,
 * tol_lin}; u, v_old, a_old updated in place.  Returns the Newton iteration
 * count, -1 if not converged, -2 det F <= 0, -3 linear solve failure.
 * [pinned: conservation of the Newmark relations checked in tests; residual
 *  at the converged state below the tolerance] */
int orc_newmark_step(int n_levels, void **meshes, void **ops,
                     const orc_mg_params *prm, const double *nm,
                     int max_newton, int min_newton, double *u, double *v_old,
                     double *a_old, int *lin_its);
/* The same with an interface: every residual adds the traction term over
 * n_faces (cell, face) pairs of the finest mesh (orc_interface_traction; t:
 * 3 values per face point).
 * [pinned: the reference's own goldens -- Newton counts and watchpoint
 *  displacements of the Turek flap, tests/test_turek.py] */
int orc_newmark_step_iface(int n_levels, void **meshes, void **ops,
                           const orc_mg_params *prm, const double *nm,
                           int max_newton, int min_newton, double *u, double *v_old,
                           double *a_old, int *lin_its, int64_t n_faces,
                           const int32_t *faces, const double *traction);
/* plain / Jacobi-preconditioned CG (omega scales the inverse diagonal) */
int orc_cg_solve(int kind, void *op, int precond, double omega, const double *b,
                 double *x, double tol_rel, int max_iter, double *final_rel);
/* Chebyshev smoother of the V-cycle levels on an explicit dense matrix A
 * (row-major n x n, symmetric positive definite, no constrained entries):
 * the same set-up as a level (eigenvalue estimate from eig_it CG steps on the
 * documented start vector; interval [1.2 lmax / range, 1.2 lmax] for
 * range > 1, or solver mode for degree <= 0 with the degree chosen from the
 * estimate and range as the target reduction), then one application:
 * zero_guess != 0: x = p(D^-1 A) D^-1 b (Chebyshev vmult); otherwise
 * x += p(D^-1 A) D^-1 (b - A x) (Chebyshev step).
 * info (optional, 5 values): lmax, lmin, theta, delta, degree.
 * [pinned: the error propagator on every eigenvector of D^-1 A equals the
 *  closed form T_k((theta-lambda)/delta)/T_k(theta/delta); the solver-mode
 *  degree is the least k with T_k(theta/delta) > 1/range;
 *  tests/test_oracle_mg.py, which also checks that planted mistakes in the
 *  recurrence fail these pins] */
void orc_cheb_dense(int n, const double *A, int degree, double range, int eig_it, int zero_guess,
                    const double *b, double *x, double *info);
/* Chebyshev smoother diagnostics on one level: returns the Lanczos max/min
 * eigenvalue estimates (of D^-1 A) from n_it CG iterations started on the
 * documented initial vector. */
void orc_cheb_eigenvalues(int kind, void *op, void *mesh, int n_it,
                          double *lmax, double *lmin);

#ifdef __cplusplus
}
#endif
