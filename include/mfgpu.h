/*
 * mfgpu: MI355X-native matrix-free operators and MG-preconditioned CG.
 *
 * C ABI of libmfgpu.so.  Every entry point returns an int status (MFG_OK = 0)
 * and never throws; mfg_last_error() returns the message of the last failure
 * on the calling thread.  Vector arguments are caller-owned DEVICE pointers
 * in the operator's precision; the library owns device copies of index maps,
 * constraint lists, geometry and quadrature caches.  All work of one object
 * is ordered on that object's HIP stream (default: the null stream), which
 * makes an object thread-compatible but not thread-safe.
 *
 * Reference interfaces replaced (davidscn/matrix-free-dealii-precice):
 *   NeoHookOperator<dim,Number>    include/solid_mechanics/mf_nh_operator.h:99-237
 *   LaplaceOperator<dim,number>    include/heat_transfer/heat_transfer.h:81-127
 *   MatrixFree data layout build   include/solid_mechanics/mf_elasticity.h:725-1055
 *   MGTransferMatrixFree / PreconditionChebyshev / SolverCG / Multigrid as
 *   configured at mf_elasticity.h:1140-1279, 1981-2070, heat_transfer.h:757-812
 */
#ifndef MFGPU_H
#define MFGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- status --- */
#define MFG_OK 0
#define MFG_ERR_ARGUMENT 1   /* invalid argument or unsupported configuration */
#define MFG_ERR_HIP 2        /* a HIP runtime call failed */
#define MFG_ERR_NUMERIC 3    /* det F <= 0, non-positive diagonal, breakdown */
#define MFG_ERR_NOCONV 4     /* iteration cap reached (SolverControl::NoConvergence) */
#define MFG_ERR_STATE 5      /* called before the required set-up step */

#define MFG_FP64 0
#define MFG_FP32 1

/* MF caching strategies (parameter_handling.h "MF caching",
 * mf_nh_operator.h:346-523, 848-1010).  All four define the same operator:
 *   scalar_referential: reference geometry + 1 scalar per point; F rebuilt
 *                       from the cached nodal displacement at every vmult
 *   tensor2:            deformed geometry + 2 scalars + tau/J (the reference
 *                       stores 18 words/point, the data the canonical roofline
 *                       counts; formulation 1 stores them folded into 16,
 *                       or 17 on meshes with non-affine cells)
 *   tensor4:            tensor2 data + the symmetric 4th-order tangent (53 words)
 *   tensor4_ns:         reference geometry + the full 4th-order tangent (91 words)
 * Formulation 0 exists for tensor2 only (mf_nh_operator.h:706). */
#define MFG_CACHING_SCALAR_REFERENTIAL 0
#define MFG_CACHING_TENSOR2 1
#define MFG_CACHING_TENSOR4 2
#define MFG_CACHING_TENSOR4_NS 3

/* Last error message of the calling thread ("" if none). */
const char *mfg_last_error(void);
/* Library version string. */
const char *mfg_version(void);

/* ------------------------------------------------- mesh / level layout --- */
/* Host-side description of one mesh level: the data a MatrixFree object is
 * built from (mf_elasticity.h:933-992, 996-1055).
 *   cell_nodes:       n_cells*(degree+1)^3 node indices, lexicographic within
 *                     the cell (a + N*(b + N*c)); dof = n_components*node+comp
 *   cell_vertices:    n_cells*8*3 vertex coordinates, vertex v = vx+2vy+4vz;
 *                     geometry is the trilinear map (MappingQ1)
 *   constrained_dofs: sorted homogeneous Dirichlet dofs (identity rows)
 *   material_ids:     per cell, or NULL for all 0 (cell->material_id(); the
 *                     solid operator uses the inclusion material of
 *                     mfg_solid_params on cells whose id is not 0,
 *                     mf_nh_operator.h:404, 656-658; MatrixFree batches
 *                     cells by it, mf_elasticity.h:837-855)
 * Node indices must be < 2^28.  n_q_1d must equal degree+1. */
typedef struct mfg_level_desc {
  int32_t degree;
  int32_t n_q_1d;
  int32_t n_components;
  int32_t n_cells;
  int64_t n_nodes;
  const int32_t *cell_nodes;
  const double *cell_vertices;
  const int32_t *constrained_dofs;
  int64_t n_constrained;
  const int32_t *material_ids;
} mfg_level_desc;

/* Structured box mesh builder (replaces GridGenerator::subdivided_hyper_rectangle
 * + DoFHandler::distribute_dofs + make_constraints, mf_elasticity.h:659-721,
 * 1918-1974, for global refinement of a box).  face_mask bit f constrains
 * boundary face f in the colorize numbering (0:x=lo,1:x=hi,2:y=lo,3:y=hi,
 * 4:z=lo,5:z=hi); comp_mask selects the constrained components. */
typedef struct mfg_mesh_t *mfg_mesh;
int mfg_box_mesh_create(const double lo[3], const double hi[3], const int32_t n_cells[3], int32_t degree,
                        int32_t n_components, int32_t face_mask, int32_t comp_mask, mfg_mesh *out);
/* The same with a component mask per boundary face (colorize numbering as
 * above): the per-boundary-id ComponentMasks of make_constraints
 * (mf_elasticity.h:1918-1974), e.g. a flap clamped on x = lo and held in z on
 * the z faces (turek_hron.h:45-58: face_comp_mask = {7, 0, 0, 0, 4, 4}). */
int mfg_box_mesh_create_bc(const double lo[3], const double hi[3], const int32_t n_cells[3], int32_t degree,
                           int32_t n_components, const int32_t face_comp_mask[6], mfg_mesh *out);
int mfg_box_mesh_destroy(mfg_mesh mesh);
/* Material id per cell (cell order cx + n0 (cy + n1 cz), n_cells entries;
 * NULL resets all to 0), handed to operators through mfg_box_mesh_describe
 * (cell->set_material_id of the reference's cases, turek_hron.h:88-90;
 * id 0 is the first material, any other id the inclusion material). */
int mfg_box_mesh_set_material_ids(mfg_mesh mesh, const int32_t *ids);
/* Fills desc with pointers into the mesh's host arrays (valid while mesh lives). */
int mfg_box_mesh_describe(mfg_mesh mesh, mfg_level_desc *desc);
/* Node coordinates, n_nodes*3 doubles, host memory. */
int mfg_box_mesh_node_coords(mfg_mesh mesh, double *out);
/* Cells per direction. */
int mfg_box_mesh_cells(mfg_mesh mesh, int32_t n_cells[3]);
/* (cell, face) pairs of a box mesh on boundary id b: 2 d for the low and
 * 2 d + 1 for the high side in direction d (deal.II's colorized box ids), the
 * local face number being b itself; faces == NULL returns the count only. */
int mfg_box_mesh_boundary_faces(mfg_mesh mesh, int32_t boundary_id, int32_t *faces, int64_t *n_faces);
/* Node numbering of a box mesh (the vector layout in HBM).  A new mesh is
 * LEXICOGRAPHIC.  CELLWISE numbers nodes cell by cell, as
 * DoFHandler::distribute_dofs does (mf_elasticity.h:667, heat_transfer.h:437;
 * the reference's Cuthill-McKee pass, mf_elasticity.h:669, is not
 * reproduced): cell c owns the block [c p^3, (c+1) p^3) of the nodes on its
 * high side, and the three low boundary planes come last.  The cell kernels'
 * gathers then touch about a fifth fewer 128-byte lines.  Renumbering permutes
 * cell_nodes, node coordinates and the Dirichlet list in place; call it before
 * describing the mesh to operators.  Operator results on the two numberings
 * are the same vectors up to the permutation. */
#define MFG_NUMBERING_LEXICOGRAPHIC 0
#define MFG_NUMBERING_CELLWISE 1
int mfg_box_mesh_renumber(mfg_mesh mesh, int32_t numbering);
int32_t mfg_box_mesh_numbering(mfg_mesh mesh);

/* ------------------------------------------------------------ operators --- */
typedef struct mfg_operator_t *mfg_operator;

/* Neo-Hookean material (material.h:78-94) and Newmark factor alpha_1 =
 * 1/(beta dt^2) (mf_elasticity.h:240); formulation 0 or 1; caching one of
 * MFG_CACHING_*.  Replaces NeoHookOperator::initialize + set_material
 * (mf_nh_operator.h:115-128, 331-388).
 * inclusion_mu, inclusion_nu, inclusion_rho: set_material's second material,
 * used on the cells whose material id is not 0 (material_inclusion,
 * mf_nh_operator.h:121-128, 529-539; the reference builds it with 100 mu,
 * mf_elasticity.h:470-476), same alpha_1 and formulation.  inclusion_mu = 0:
 * the inclusion cells use the first material. */
typedef struct mfg_solid_params {
  double mu, nu, rho, alpha_1;
  int32_t formulation;
  int32_t caching;
  double inclusion_mu, inclusion_nu, inclusion_rho;
} mfg_solid_params;

int mfg_solid_create(const mfg_level_desc *level, const mfg_solid_params *params, int32_t precision,
                     mfg_operator *out);
/* LaplaceOperator (heat_transfer.h:81-250): A = M + dt*c(x)*K.  coefficient is
 * a host table per (cell, q) or NULL for c == 1 (Coefficient, :52-78). */
int mfg_heat_create(const mfg_level_desc *level, double dt, const double *coefficient, int32_t precision,
                    mfg_operator *out);
/* Physical coordinates of an operator's quadrature points (any kind):
 * n_cells * (p+1)^3 * 3 doubles into device memory, input cell order, point
 * i + N (j + N k) -- where the reference evaluates source terms and exact
 * solutions (FEEvaluation::quadrature_point, heat_transfer.h:610-611, 836). */
int mfg_operator_quadrature_points(mfg_operator op, double *xyz);
/* LaplaceProblem::assemble_rhs (heat_transfer.h:577-655) without the preCICE
 * interface flux: rhs = int (u_old + dt f) phi dx - A u_D, where A is the
 * heat operator without constraints (the reference's inhomogeneous_operator)
 * and u_D holds the Dirichlet values on the constrained dofs and zero
 * elsewhere (constraints.distribute of a zero vector, :587-588).  u_old and
 * u_D are read as stored.  f: the source at the quadrature points (device,
 * layout of mfg_operator_quadrature_points without the 3; NULL = 0).  The
 * constrained rows of rhs are set to zero (the solve's homogeneous
 * constraints).  FP64 operator; device vectors; rhs must not alias an input. */
int mfg_heat_assemble_rhs(mfg_operator op, double *rhs, const double *u_old, const double *u_dirichlet,
                          const double *f_qp);
/* LaplaceProblem::compute_error (heat_transfer.h:815-848):
 * sqrt(sum over cells and points of (exact - u_h)^2 JxW), u_h read as stored,
 * exact at the quadrature points (device, layout as f_qp).  Summed in a fixed
 * order.  FP64 operator. */
int mfg_heat_l2_error(mfg_operator op, const double *u, const double *exact_qp, double *error);
int mfg_operator_destroy(mfg_operator op);                         /* clear() :112-113 */
int mfg_operator_set_stream(mfg_operator op, void *hip_stream);
int64_t mfg_operator_m(mfg_operator op);                             /* m() :133 */
int64_t mfg_operator_n(mfg_operator op);                             /* n() :135 */
int32_t mfg_operator_precision(mfg_operator op);

/* cache(): linearise at the device displacement u (precision u_precision),
 * NeoHookOperator::cache (mf_nh_operator.h:392-523).  MFG_ERR_NUMERIC if
 * det F <= 0 at any quadrature point. */
int mfg_solid_cache(mfg_operator op, const void *u, int32_t u_precision);

/* vmult / vmult_add / Tvmult / Tvmult_add (mf_nh_operator.h:543-625,
 * MatrixFreeOperators::Base::vmult for heat): dst = A src (dst += A src),
 * identity rows on constrained dofs.  Tvmult == vmult (A is symmetric). */
int mfg_operator_vmult(mfg_operator op, void *dst, const void *src);
int mfg_operator_vmult_add(mfg_operator op, void *dst, const void *src);
int mfg_operator_Tvmult(mfg_operator op, void *dst, const void *src);
int mfg_operator_Tvmult_add(mfg_operator op, void *dst, const void *src);

/* compute_diagonal (mf_nh_operator.h:1023-1060, heat_transfer.h:205-230):
 * diagonal with constrained entries set to 1, and its inverse. */
int mfg_operator_compute_diagonal(mfg_operator op);
/* Copies (device to device) the diagonal / inverse diagonal; either may be NULL.
 * get_matrix_diagonal_inverse (:157-161). */
int mfg_operator_copy_diagonal(mfg_operator op, void *diag, void *inv_diag);
/* el(row, col): diagonal entry only (:1064-1073). */
int mfg_operator_el(mfg_operator op, int64_t row, int64_t col, double *value);
/* precondition_Jacobi: dst = omega * D^{-1} src (:281-291). */
int mfg_operator_precondition_jacobi(mfg_operator op, void *dst, const void *src, double omega);
/* memory_consumption (:241-260): bytes held for vmult (caches + layout). */
size_t mfg_operator_memory_consumption(mfg_operator op);
/* Quadrature-point cache as stored on the device (structure of arrays,
 * field-major: field f of point (cell, q) at f*n_cells*nq3 + cell*nq3 + q).
 * Solid tensor2 formulation 1: Jinv_cur[f][e] (9), JxW_ref s, JxW_ref tau
 * (xx,yy,zz,xy,xz,yz), with s = mu - 2 lambda ln J and JxW_ref = JxW_cur / J
 * (the reference's JxW_cur 2s/J / 2 and JxW_cur tau/J); a 17th field JxW_ref
 * when some cell is not affine in the reference configuration (otherwise it is
 * a per-cell determinant times the quadrature weight).  Heat: 11 fields
 * Jinv (9), JxW, c.  Copies n_fields*n_qp values to the device pointer out;
 * n_fields may be queried with out == NULL. */
int mfg_operator_qp_data(mfg_operator op, void *out, int32_t *n_fields);
/* Point storage of the tensor2, formulation-1 cache (no counterpart in the
 * reference, whose storage is fixed; the operator is the same either way).
 * MFG_T2_FOLDED: Jinv_cur (9), JxW_ref s, JxW_ref tau (6), plus JxW_ref on
 * meshes with a cell that is not affine -- the layout mfg_operator_qp_data
 * documents.  MFG_T2_GRADIENT (meshes whose cells are all affine): the
 * reference-cell gradient of the linearisation point Grad_xi u[d][f] (9) and
 * s = mu - 2 lambda ln J, 10 words per point; vmult rebuilds Jinv_cur =
 * (J_ref + Grad_xi u)^{-1} and tau = mu F F^T - s I per point from the
 * per-cell J_ref.  Default: GRADIENT on meshes whose cells are all affine,
 * FOLDED otherwise.  Changing the storage discards the cache: call
 * mfg_operator_cache (and compute_diagonal) again. */
#define MFG_T2_FOLDED 0
#define MFG_T2_GRADIENT 1
int mfg_operator_set_point_storage(mfg_operator op, int32_t storage);
int32_t mfg_operator_point_storage(mfg_operator op);
/* Coupling interface of the solid (Solid::assemble_residual,
 * mf_elasticity.h:1849-1897): rhs += int t0 . w dA over the listed faces,
 * with the spatial traction t pulled back at the displacement u as
 * t0 = |det F F^{-T} N| t.  faces: host array of n_faces pairs (input cell
 * index, local face 2 axis + side, side 1 = the cell's high side in reference
 * direction axis).  traction: device pointer, 3 values per face point, point
 * q = qa + (p+1) qb over the Gauss points of the face's two tangential
 * reference directions in ascending order (what preCICE's read gives per
 * quadrature point, mf_elasticity.h:1876-1878).  u, rhs: FP64 device vectors
 * of the operator's layout; constrained dofs are left as they are (the
 * reference zeroes them afterwards, :1907).  FP64 solid operators only.
 * Faces share edge nodes: the sums use atomics (order-dependent rounding). */
int mfg_solid_add_interface_traction(mfg_operator op, int64_t n_faces, const int32_t *faces, const double *u,
                                     const double *traction, double *rhs);
/* Reference-configuration coordinates of the same face points (the coupling
 * mesh the reference writes to preCICE, include/adapter/quad_interface.h:
 * 198-340): 3 doubles per point into the device pointer xyz. */
int mfg_operator_face_quadrature_points(mfg_operator op, int64_t n_faces, const int32_t *faces, double *xyz);
/* Values of an FP64 field u of the operator's layout (n_components per node)
 * at the same face points: n_components doubles per point into the device
 * pointer values -- the data the reference writes to preCICE at the interface
 * quadrature points (displacement, quad_interface.h:198-340; temperature). */
int mfg_operator_face_values(mfg_operator op, int64_t n_faces, const int32_t *faces, const double *u,
                             double *values);
/* Heat interface flux (LaplaceProblem::assemble_rhs, heat_transfer.h:618-652):
 * rhs -= dt int q phi dA over the faces, q = flux per face point (device
 * pointer, point order as above), dt the operator's; constrained dofs are left
 * as they are.  FP64 rhs; atomics on shared edge nodes. */
int mfg_heat_add_interface_flux(mfg_operator op, int64_t n_faces, const int32_t *faces, const double *flux,
                                double *rhs);
/* Number of cell colours of the device layout.  Cells are stored
 * colour-major; cells of one colour share no node, so each colour is one
 * launch whose scatter needs no atomics (deterministic results).  Replaces the
 * role of MatrixFree's cell-batch partition (mf_elasticity.h:806-859). */
int32_t mfg_operator_n_colours(mfg_operator op);

/* Cell schedule of vmult.  MFG_CELLS_COLOURS (the default above 2^20
 * quadrature points): one launch per cell colour, a workgroup per group of
 * cells.  MFG_CELLS_GATHER: one
 * launch over all cells storing each point's contribution, then one launch
 * summing every node's contributions in colour order -- bitwise the COLOURS
 * result, in two launches instead of one per colour (the default up to 2^20
 * quadrature points, where colour launches are latency bound); it holds
 * N_qp * n_components extra values. */
#define MFG_CELLS_COLOURS 0
#define MFG_CELLS_GATHER 3
int mfg_operator_set_cell_schedule(mfg_operator op, int32_t schedule);
int32_t mfg_operator_cell_schedule(mfg_operator op);

/* Kernel timing (TimerOutput-style sections of the reference,
 * mf_elasticity.h:427-432): when enabled, every vmult/vmult_add records HIP
 * events on the operator's stream around its cell kernel.  mfg_operator_timing
 * waits for the recorded events, returns the summed cell-kernel time and the
 * launch count since the last query, and resets both. */
int mfg_operator_enable_timing(mfg_operator op, int32_t enable);
int mfg_operator_timing(mfg_operator op, double *total_ms, int64_t *launches);

/* ------------------------------------------------------ multigrid transfer --- */
/* Transfer between nested levels (MGTransferMatrixFree, used at
 * mf_elasticity.h:863-877, 920-922; heat_transfer.h:761-762).  children holds,
 * for every coarse cell (input order), its 8 fine cells (input order) with
 * child index dx + 2 dy + 4 dz.  Operators may differ in precision.  */
typedef struct mfg_transfer_t *mfg_transfer;
int mfg_transfer_create(mfg_operator coarse, mfg_operator fine, const int32_t *children, mfg_transfer *out);
int mfg_transfer_destroy(mfg_transfer t);
/* children table of a box mesh and its once-refined box (n_fine = 2 n_coarse) */
int mfg_box_mesh_children(mfg_mesh coarse, mfg_mesh fine, int32_t *children);
/* fine += P coarse (prolongate_and_add); coarse constrained read as zero,
 * fine constrained dofs untouched.  Vectors in the operators' precisions. */
int mfg_transfer_prolongate_add(mfg_transfer t, void *fine, const void *coarse);
/* coarse += P^T fine (restrict_and_add); fine constrained read as zero,
 * coarse constrained untouched. */
int mfg_transfer_restrict_add(mfg_transfer t, void *coarse, const void *fine);
/* coarse = nodal interpolant of the fine function (interpolate_to_mg). */
int mfg_transfer_interpolate(mfg_transfer t, void *coarse, const void *fine);

/* ------------------------------------------------ multigrid preconditioner --- */
/* Geometric V-cycle with Chebyshev smoothing (Multigrid + PreconditionMG +
 * mg::SmootherRelaxation<PreconditionChebyshev>, mf_elasticity.h:1140-1279,
 * heat_transfer.h:757-812).  Levels > 0: Chebyshev of `cheb_degree` with the
 * inverse diagonal, interval [1.2 lmax/smoothing_range, 1.2 lmax], lmax from
 * eig_cg_iterations Jacobi-preconditioned CG steps.  Level 0: coarse_kind 0 =
 * Chebyshev as a solver (range coarse_range, degree from the estimate, m()
 * eigenvalue iterations, or min(m(), coarse_eig_iterations) when that is
 * > 0); 1 = CG to relative coarse_tolerance (absolute 1e-10, max(m,100)
 * steps) preconditioned by the level-0 Chebyshev. */
typedef struct mfg_mg_params {
  int32_t cheb_degree;
  double smoothing_range;
  int32_t eig_cg_iterations;
  int32_t coarse_kind;
  double coarse_range;
  double coarse_tolerance;
  int32_t coarse_eig_iterations; /* 0: m() (the reference's choice) */
} mfg_mg_params;
typedef struct mfg_mg_t *mfg_mg;
/* ops[0..n_levels-1] coarse to fine; transfers[l-1] connects level l-1 and l.
 * The hierarchy references (does not own) operators and transfers. */
int mfg_mg_create(int32_t n_levels, const mfg_operator *ops, const mfg_transfer *transfers,
                  const mfg_mg_params *params, mfg_mg *out);
int mfg_mg_destroy(mfg_mg mg);
/* (Re-)initialise after the level operators changed (cache()): diagonals,
 * eigenvalue estimates, Chebyshev coefficients (setup_gmg, :1140-1279). */
int mfg_mg_setup(mfg_mg mg);
/* dst = V-cycle(src) (PreconditionMG::vmult); vec_precision of dst/src. */
int mfg_mg_vmult(mfg_mg mg, void *dst, const void *src, int32_t vec_precision);
/* Chebyshev data of a level: lmax, lmin estimates and polynomial degree. */
int mfg_mg_level_info(mfg_mg mg, int32_t level, double *lmax, double *lmin, int32_t *degree);
/* V-cycle launch mode (no counterpart in the reference, whose V-cycle is a
 * host loop over levels).  enable != 0: from the second V-cycle on, the
 * sweeps above the coarse level replay as two captured hipGraphs on a stream
 * of the hierarchy's own, ordered after and before the fine operator's
 * stream; the coarse solve runs eagerly between them.  Results are bitwise
 * those of the eager V-cycle.  0 (the default): every V-cycle launches
 * eagerly -- on MI355X the graphs are not faster (the small levels are bound
 * by each colour launch's critical path) and re-capturing after every setup
 * adds latency.
 * graph_nodes (nullable) receives the node count of the captured graphs. */
int mfg_mg_use_graph(mfg_mg mg, int32_t enable);
int mfg_mg_graph_info(mfg_mg mg, int64_t *graph_nodes);

/* ------------------------------------------------------------------- CG --- */
/* SolverCG (Solid::solve_linear_system, mf_elasticity.h:1981-2070): solve
 * A x = b from the initial x to ||r|| <= tol_rel*||b||.  preconditioner:
 * MFG_PRECOND_NONE, MFG_PRECOND_JACOBI (omega * D^-1, PreconditionJacobi) or
 * MFG_PRECOND_MG (mg must be set).  Vectors in A's precision.  Returns
 * MFG_ERR_NOCONV when max_iter is reached (info still filled). */
#define MFG_PRECOND_NONE 0
#define MFG_PRECOND_JACOBI 1
#define MFG_PRECOND_MG 2
typedef struct mfg_solve_info {
  int32_t iterations;
  double initial_residual;
  double final_residual;
  double condition_estimate;  /* from the Lanczos coefficients of the solve */
} mfg_solve_info;
int mfg_cg_solve(mfg_operator A, int32_t preconditioner, mfg_mg mg, double omega, const void *b, void *x,
                 double tol_rel, int32_t max_iter, mfg_solve_info *info);
/* mfg_cg_solve without device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The CG vectors, reduction
 * partials and step-length scalars live in a file-scope device array of
 * 32 MiB + 8 KiB that each GPU the library runs on holds until the process
 * ends; the first call on a device looks up its address (and loads the module
 * holding it).  Limit: operators of at most 2^20 rows (MFG_ERR_ARGUMENT
 * beyond); the multigrid preconditioner must not replay graphs
 * (mfg_mg_use_graph off).  Every call on one GPU shares that array: calls
 * that overlap on one GPU race.  Same arguments, results and errors as
 * mfg_cg_solve. */
int mfg_cg_solve_noalloc(mfg_operator A, int32_t preconditioner, mfg_mg mg, double omega, const void *b, void *x,
                         double tol_rel, int32_t max_iter, mfg_solve_info *info);

/* ------------------------------------------------ Newton-Newmark driver --- */
/* Residual of the solid at displacement u and acceleration acc, on the
 * operator's level layout (Solid::assemble_residual fast path,
 * mf_elasticity.h:1764-1846, then constraints.set_zero, :1905): R (FP64)
 * = -int tau F^-T : Grad w + int rho (body - acc) . w; returns ||R||_2.
 * The operator must be FP64. */
int mfg_solid_residual(mfg_operator op, const double *u, const double *acc, const double body[3], double *R,
                       double *norm);

/* Newmark beta/gamma time stepping with full Newton-Raphson and MG-CG linear
 * solves (Solid::solve_nonlinear_timestep, update_acceleration/velocity,
 * check_convergence, solve_linear_system; mf_elasticity.h:1357-1491,
 * 1600-1620, 1981-2070).  Convergence is tested from Newton iteration
 * min_newton on (the reference uses 3): ||R|| <= tol_f_abs or ||du|| <=
 * tol_u_abs, or ||R||/||R_0|| <= tol_f and ||du||/||du_0|| <= tol_u. */
typedef struct mfg_newmark_params {
  double beta, gamma, dt;
  double body_force[3];
  int32_t max_newton;
  int32_t min_newton;
  double tol_f, tol_u, tol_f_abs, tol_u_abs;
  double tol_lin;   /* linear solve: ||r|| <= tol_lin ||R|| */
  int32_t max_lin;  /* CG iteration cap per linear solve */
  /* Condition estimate before every linear solve: this many plain CG
   * iterations on the fine operator (stopping early at ||r|| <= tol_lin ||R||;
   * with FP32 levels on the FP32 fine-level operator), their Lanczos condition
   * number reported and the iterate discarded
   * (solve_linear_system, mf_elasticity.h:1998-2013; the reference's default
   * is on with 20 iterations, parameter_handling.h:183-184).  0 skips it. */
  int32_t cond_cg_iterations;
  /* Preconditioner of the linear solves ("Preconditioner type",
   * Solid::solve_linear_system, mf_elasticity.h:2026-2051): MFG_NEWTON_GMG
   * (0, the default), MFG_NEWTON_JACOBI (PreconditionJacobi with
   * jacobi_relaxation, the reference's "Preconditioner relaxation") or
   * MFG_NEWTON_NONE.  As in update_matrix_free (:1320-1324) the multigrid
   * levels are only set up for MFG_NEWTON_GMG. */
  int32_t linear_preconditioner;
  double jacobi_relaxation;
} mfg_newmark_params;
#define MFG_NEWTON_GMG 0
#define MFG_NEWTON_JACOBI 1
#define MFG_NEWTON_NONE 2
typedef struct mfg_step_info {
  int32_t newton_iterations;
  int32_t linear_iterations;  /* summed over the Newton iterations */
  int32_t converged;
  double residual_0, residual_final, update_final;
  double seconds;
  double condition_estimate;  /* of the last linear solve (1 when not estimated) */
} mfg_step_info;
typedef struct mfg_solid_problem_t *mfg_solid_problem;
/* Level hierarchy coarse..fine (levels[n_levels-1] is the active mesh);
 * children[l-1] connects level l-1 and l (see mfg_transfer_create).
 * level_precision: MFG_FP64 (plain FP64 V-cycle) or MFG_FP32 (FP32 V-cycle
 * and level caches; the coarse level, outer CG, residual and Newton update
 * stay FP64).  material->alpha_1 is ignored (1/(beta dt^2) is used);
 * material->caching selects the operators' caching (initialize's `caching`
 * argument, mf_nh_operator.h:115-119).  The tensor2 operators of
 * formulation 1 on all-affine meshes (FP64 and FP32) keep the
 * MFG_T2_GRADIENT point storage. */
int mfg_solid_problem_create(int32_t n_levels, const mfg_level_desc *levels, const int32_t *const *children,
                             const mfg_solid_params *material, const mfg_newmark_params *newmark,
                             const mfg_mg_params *mg, int32_t level_precision, mfg_solid_problem *out);
int mfg_solid_problem_destroy(mfg_solid_problem p);
int mfg_solid_problem_set_stream(mfg_solid_problem p, void *hip_stream);
/* Device FP64 state: total displacement u (linearisation point and
 * old_displacement), velocity and acceleration of the last converged step.
 * Any pointer may be NULL (left unchanged / not copied). */
int mfg_solid_problem_set_state(mfg_solid_problem p, const double *u, const double *v, const double *a);
int mfg_solid_problem_get_state(mfg_solid_problem p, double *u, double *v, double *a);
/* One Newmark time step (Newton to convergence); MFG_ERR_NOCONV if
 * max_newton iterations do not converge (the state is then left at the last
 * iterate). */
int mfg_solid_problem_step(mfg_solid_problem p, mfg_step_info *info);
/* Device bytes held by the problem (operators, caches, MG, state vectors). */
/* Coupling interface of the time step: every residual of the following
 * steps adds the traction term over the faces, pulled back at the current
 * displacement (as mfg_solid_add_interface_traction; mf_elasticity.h:
 * 1849-1897).  faces: host array of n_faces (input cell, local face) pairs of
 * the finest level; traction: device pointer to 3 values per face point,
 * owned by the caller, read at every residual -- the caller refreshes it
 * between steps (preCICE read, mf_elasticity.h:1876-1878) and keeps it alive
 * while steps run.  n_faces = 0 removes the interface. */
int mfg_solid_problem_set_interface(mfg_solid_problem p, int64_t n_faces, const int32_t *faces,
                                    const double *traction);
size_t mfg_solid_problem_memory(mfg_solid_problem p);

/* ---------------------------------------------------------- serving loop --- */
/* Continuous-batching service of Newmark time steps for many live solid
 * sessions (no counterpart in the reference, which steps one solid per MPI
 * job; each step is Solid::solve_nonlinear_timestep).  Session state (u, v,
 * a, caches, MG hierarchy) stays on the device while resident.  Requests are
 * released at their arrival times; up to max_concurrent steps run at once,
 * each on its own HIP stream and host thread, so their work overlaps on the
 * device.  A request's session is admitted when its estimated device bytes
 * fit hbm_budget_bytes; idle sessions are evicted least-recently-used (state
 * saved to host) to make room.  Requests of one session run in arrival
 * order.  Results are deterministic: a session's steps give identical
 * numbers alone and in the mix. */
typedef struct mfg_session_spec {
  double lo[3], hi[3];
  int32_t n_cells[3];         /* finest level; coarser levels halve while all counts stay even */
  int32_t degree;
  mfg_solid_params material;  /* alpha_1 ignored */
  mfg_newmark_params newmark;
  mfg_mg_params mg;
  int32_t level_precision;    /* MFG_FP64 or MFG_FP32 (mixed) */
  double u0_amplitude, u0_noise; /* initial u: amp sin sin sin + noise U[-1,1] */
  uint64_t seed;
} mfg_session_spec;
typedef struct mfg_request {
  int32_t session;
  double arrival_s;  /* seconds after the start of the trace */
} mfg_request;
typedef struct mfg_request_result {
  double arrival_s, start_s, end_s;
  int32_t newton_iterations, linear_iterations, status;
  double residual_final;
} mfg_request_result;
typedef struct mfg_server_params {
  int32_t max_concurrent;
  size_t hbm_budget_bytes;
  int32_t real_time;  /* 1: honour arrival times; 0: all requests available at t = 0 (latency from t = 0) */
} mfg_server_params;
typedef struct mfg_serve_stats {
  double makespan_s, steps_per_s, latency_p50_s, latency_p99_s, latency_mean_s;
  int32_t completed, failed, admissions, evictions, max_resident;
  size_t peak_bytes;
} mfg_serve_stats;
/* Runs the request trace.  results: n_requests entries (request order).
 * final_u: optional array of n_sessions host pointers receiving each
 * session's final displacement (n_dofs doubles; entries may be NULL), in the
 * CELLWISE numbering of the session's box (every level of a session is
 * renumbered so; mfg_box_mesh_renumber gives the node coordinates).  */
int mfg_serve(const mfg_session_spec *sessions, int32_t n_sessions, const mfg_request *requests, int32_t n_requests,
              const mfg_server_params *params, mfg_request_result *results, double *const *final_u,
              mfg_serve_stats *stats);
/* DoF count of a session's finest level (to size final_u buffers). */
int64_t mfg_session_n_dofs(const mfg_session_spec *spec);

/* ------------------------------------------------------ batched instances --- */
/* Many independent operators (heat and solid, degrees 1..6, uneven meshes)
 * applied in one launch (no counterpart in the reference, which serves one
 * operator per MatrixFree object; each instance's arithmetic is that of
 * NeoHookOperator::vmult, mf_nh_operator.h:543-625, or LaplaceOperator,
 * heat_transfer.h:81-250, and batched results equal per-instance results).
 * Cell work is cut into tasks of up to cells_per_task cells of one colour of
 * one instance.  Schedules: MFG_SCHEDULE_QUEUE, one persistent launch per
 * application (workgroups take tasks from an atomic ticket; a task waits for
 * its instance's previous colour), MFG_SCHEDULE_COLOUR, one launch per
 * colour, MFG_SCHEDULE_STREAMS, every instance's own operator launches
 * spread over several streams (instances assigned by work, longest first), so
 * that small instances overlap, or MFG_SCHEDULE_GATHER (batched vmult only):
 * the instances of one (kind, degree) packed into units of similar work, each
 * unit one launch over the cells of all its instances into their
 * contribution buffers and one launch of the ordered per-node sums, units
 * spread over several streams (each instance's GATHER tables and buffer are
 * its operator's: a batch and a GATHER application of one of its operators on
 * another stream must not overlap).  The tuner times them and the cells per task of every
 * (kind, degree) present, and keeps the fastest.  For the batched solver,
 * STREAMS runs every instance's own mfg_cg_solve / mfg_mg_vmult on worker
 * threads with a stream each. */
#define MFG_SCHEDULE_COLOUR 0
#define MFG_SCHEDULE_QUEUE 1
#define MFG_SCHEDULE_STREAMS 2
#define MFG_SCHEDULE_GATHER 3
#define MFG_BATCH_SOLID_F1 0 /* tensor2, formulation 1 */
#define MFG_BATCH_SOLID_F0 1 /* tensor2, formulation 0 */
#define MFG_BATCH_HEAT 2
#define MFG_BATCH_SOLID_F1_GRADIENT 3 /* tensor2, formulation 1, MFG_T2_GRADIENT point storage */
typedef struct mfg_batch_tuning {
  int32_t schedule;
  int32_t cells_per_task[4][7]; /* [kind][degree] */
  int32_t waves_per_simd;       /* register budget of the kernel (2 or 4) */
  double ms_colour, ms_queue;   /* one application of the whole batch */
  double ms_streams;            /* the same with MFG_SCHEDULE_STREAMS (batched solver: one
                                   preconditioner application of all instances) */
  double ms_gather;             /* the same with MFG_SCHEDULE_GATHER (batched vmult) */
} mfg_batch_tuning;
typedef struct mfg_batch_t *mfg_batch;
/* All operators must be cached and share one precision.  The batch holds the
 * operators' device pointers: an operator whose device buffers are
 * reallocated later (mfg_operator_set_point_storage, say) makes the next
 * application rebuild the batch's tables, and one that is not cached when the
 * batch is applied is refused (MFG_ERR_STATE). */
int mfg_batch_create(int32_t n, const mfg_operator *ops, mfg_batch *out);
int mfg_batch_destroy(mfg_batch b);
int mfg_batch_set_stream(mfg_batch b, void *hip_stream);
int mfg_batch_set_schedule(mfg_batch b, int32_t schedule);
/* cells = 0 restores the default (as many as fit 512 threads). */
int mfg_batch_set_cells_per_task(mfg_batch b, int32_t kind, int32_t degree, int32_t cells);
/* dst[i] = A_i src[i] (vmult) / dst[i] += A_i src[i] (vmult_add) for all i. */
int mfg_batch_vmult(mfg_batch b, void *const *dst, const void *const *src);
int mfg_batch_vmult_add(mfg_batch b, void *const *dst, const void *const *src);
/* Waits for the batch's stream and reports a kernel wait timeout. */
int mfg_batch_synchronize(mfg_batch b);
/* Autotuner (times vmult on the given vectors); result may be NULL. */
int mfg_batch_tune(mfg_batch b, const void *const *src, void *const *dst, int32_t repetitions,
                   mfg_batch_tuning *result);
int mfg_batch_get_tuning(mfg_batch b, mfg_batch_tuning *result);

/* MG-preconditioned CG on many instances in lockstep (the batched form of
 * mfg_cg_solve with MFG_PRECOND_MG).  fine[i] is instance i's operator
 * (one precision for all, normally FP64); mgs[i] its set-up hierarchy whose
 * finest level has the same mesh (levels may be FP32) and whose coarse
 * solver is Chebyshev (coarse_kind 0).  The V-cycles of all instances run
 * level by level (levels aligned at the coarse end) as one fixed program of
 * batched launches, captured once into a hipGraph and replayed (use_graph,
 * default on).  Every instance performs the operation sequence of its own
 * mfg_mg_vmult / mfg_cg_solve.  Its programs hold the operators' device
 * pointers: after an operator's buffers are reallocated (a new point storage,
 * say) the solver refuses to run (MFG_ERR_STATE) and must be created again. */
typedef struct mfg_batch_solver_t *mfg_batch_solver;
int mfg_batch_solver_create(int32_t n, const mfg_operator *fine, const mfg_mg *mgs, mfg_batch_solver *out);
int mfg_batch_solver_destroy(mfg_batch_solver s);
int mfg_batch_solver_use_graph(mfg_batch_solver s, int32_t on);
int mfg_batch_solver_set_schedule(mfg_batch_solver s, int32_t schedule);
/* Tunes the fine-operator and level-operator launches; outputs may be NULL. */
int mfg_batch_solver_tune(mfg_batch_solver s, int32_t repetitions, mfg_batch_tuning *fine,
                          mfg_batch_tuning *levels);
/* dst[i] = V-cycle_i(src[i]) (fine precision vectors). */
int mfg_batch_solver_precondition(mfg_batch_solver s, void *const *dst, const void *const *src);
/* Solves A_i x_i = b_i to ||r_i|| <= tol_rel ||b_i||; x holds the initial
 * guesses.  status[i]: MFG_OK, MFG_ERR_NOCONV or MFG_ERR_NUMERIC (breakdown). */
int mfg_batch_solver_solve(mfg_batch_solver s, const void *const *b, void *const *x, double tol_rel,
                           int32_t max_iter, mfg_solve_info *info, int32_t *status);
/* Launches per V-cycle program and nodes of the captured graph (0 before
 * the first graph replay). */
int mfg_batch_solver_info(mfg_batch_solver s, int32_t *vcycle_launches, int32_t *graph_nodes);

/* ------------------------------------------------------ synthetic inputs --- */
/* dst[i] = 2*U(seed,i) - 1 with U = (splitmix64(seed,i) >> 11) * 2^-53, on the
 * device (precision MFG_FP64/MFG_FP32), ordered on `stream`. */
int mfg_fill_uniform(uint64_t seed, int64_t n, int32_t precision, void *dst, void *stream);
/* Linearisation point on the host: nodal interpolant of
 * amp*sin(pi x)sin(pi y)sin(pi z) per component plus noise*U[-1,1] (seeded as
 * mfg_fill_uniform, indexed by the lexicographic dof, so the field is the same
 * under any numbering), constrained dofs zero. */
int mfg_box_mesh_fill_u0(mfg_mesh mesh, double amp, double noise, uint64_t seed, double *out);

#ifdef __cplusplus
}
#endif
#endif /* MFGPU_H */
