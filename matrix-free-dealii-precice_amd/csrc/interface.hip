// Coupling interface of the solid: the traction term of the residual on the
// interface faces (Solid::assemble_residual, mf_elasticity.h:1849-1897) and
// the coordinates of the face quadrature points that the reference hands to
// preCICE as its coupling mesh (include/adapter/quad_interface.h:198-340).
// The preCICE calls themselves stay with the caller on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "common.hpp"
#include "operator.hpp"
#include "setup_kernels.hpp"

namespace mfg {

constexpr int kFaceMaxN = kMaxDegree + 1;

// 1D tables of the GLL Lagrange basis, passed by value: values and
// derivatives at the N Gauss points of a face's tangential directions
// (val[q * N + a], der[q * N + a]), derivatives at the end points 0 and 1 in
// the normal direction (der_end[side * N + a]; the values there are Kronecker
// deltas), the Gauss points and weights on [0, 1].
struct FaceTables {
  int N;
  double val[kFaceMaxN * kFaceMaxN], der[kFaceMaxN * kFaceMaxN], der_end[2 * kFaceMaxN];
  double points[kFaceMaxN], weights[kFaceMaxN];
};

static FaceTables face_tables(int p) {
  FaceTables t{};
  const int N = p + 1;
  t.N = N;
  const std::vector<double> z = gauss_lobatto_points(p);
  std::vector<double> q, w;
  gauss_points(N, q, w);
  std::vector<double> v(N), d(N);
  for (int i = 0; i < N; ++i) {
    lagrange_basis(z, q[i], v.data(), d.data());
    for (int a = 0; a < N; ++a) {
      t.val[i * N + a] = v[a];
      t.der[i * N + a] = d[a];
    }
    t.points[i] = q[i];
    t.weights[i] = w[i];
  }
  for (int side = 0; side < 2; ++side) {
    lagrange_basis(z, (double)side, v.data(), d.data());
    for (int a = 0; a < N; ++a) t.der_end[side * N + a] = d[a];
  }
  return t;
}

// local face f = 2 axis + side; the tangential axes in ascending order
__host__ __device__ inline void face_axes(int face, int &axis, int &side, int &ta, int &tb) {
  axis = face >> 1;
  side = face & 1;
  ta = axis == 0 ? 1 : 0;
  tb = axis == 2 ? 1 : 2;
}

struct TractionArgs {
  const uint32_t *cell_nodes;  // stored cell order
  const double *vertices;      // stored cell order, 24 per cell
  const int32_t *faces;        // n_faces pairs (stored cell, local face)
  const double *u, *traction;
  double *rhs;
};

// One workgroup per face, one thread per face point and then per face node.
// At a point: F = I + Grad_xi u J_ref^{-1}, N the outward unit normal of the
// reference configuration, dA = det J_ref |J_ref^{-T} N_xi| w; the pulled-back
// traction t0 = |det F F^{-T} N| t (mf_elasticity.h:1880-1891) is integrated
// against the face's nodal basis functions, which are the only ones not
// vanishing there (GLL nodes).  Faces share edge nodes, so the sum into rhs is
// atomic.
__global__ void __launch_bounds__(64) interface_traction_kernel(TractionArgs a, FaceTables tb) {
  const int N = tb.N, N2 = N * N, N3 = N2 * N;
  __shared__ double s_u[3 * kFaceMaxN * kFaceMaxN * kFaceMaxN];
  __shared__ double s_t[3 * kFaceMaxN * kFaceMaxN];
  const int f = blockIdx.x, t = threadIdx.x;
  const int cell = a.faces[2 * f], face = a.faces[2 * f + 1];
  int axis, side, ta, tbx;
  face_axes(face, axis, side, ta, tbx);
  const uint32_t *map = a.cell_nodes + (int64_t)cell * N3;
  for (int x = t; x < N3; x += blockDim.x) {
    const int64_t node = map[x] & kNodeMaskDev;
    for (int c = 0; c < 3; ++c) s_u[x * 3 + c] = a.u[node * 3 + c];  // read plain, like read_dof_values_plain
  }
  __syncthreads();
  if (t < N2) {
    const int qa = t % N, qb = t / N;
    // per direction: basis values and derivatives at this point's coordinate
    double v[3][kFaceMaxN], d[3][kFaceMaxN];
    for (int e = 0; e < N; ++e) {
      v[ta][e] = tb.val[qa * N + e];
      d[ta][e] = tb.der[qa * N + e];
      v[tbx][e] = tb.val[qb * N + e];
      d[tbx][e] = tb.der[qb * N + e];
      v[axis][e] = (side ? e == N - 1 : e == 0) ? 1.0 : 0.0;
      d[axis][e] = tb.der_end[side * N + e];
    }
    double gr[3][3] = {};
    for (int k = 0; k < N; ++k)
      for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
          const int x = i + N * (j + N * k);
          const double g0 = d[0][i] * v[1][j] * v[2][k], g1 = v[0][i] * d[1][j] * v[2][k], g2 = v[0][i] * v[1][j] * d[2][k];
          for (int c = 0; c < 3; ++c) {
            const double uc = s_u[x * 3 + c];
            gr[c][0] += uc * g0;
            gr[c][1] += uc * g1;
            gr[c][2] += uc * g2;
          }
        }
    double xi[3];
    xi[axis] = side;
    xi[ta] = tb.points[qa];
    xi[tbx] = tb.points[qb];
    double J[3][3], Ji[3][3];
    trilinear_jacobian(a.vertices + (int64_t)cell * 24, xi[0], xi[1], xi[2], J);
    inv3(J, Ji);
    const double detJ = det3(J);
    double F[3][3], Fi[3][3];
    for (int r = 0; r < 3; ++r)
      for (int e = 0; e < 3; ++e)
        F[r][e] = gr[r][0] * Ji[0][e] + gr[r][1] * Ji[1][e] + gr[r][2] * Ji[2][e] + (r == e ? 1.0 : 0.0);
    inv3(F, Fi);
    const double detF = det3(F);
    // J_ref^{-T} N_xi with N_xi = +-e_axis
    double m[3];
    for (int e = 0; e < 3; ++e) m[e] = (side ? 1.0 : -1.0) * Ji[axis][e];
    const double mn = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    double ns[3];
    for (int r = 0; r < 3; ++r) ns[r] = detF * (Fi[0][r] * m[0] + Fi[1][r] * m[1] + Fi[2][r] * m[2]) / mn;
    const double nsn = sqrt(ns[0] * ns[0] + ns[1] * ns[1] + ns[2] * ns[2]);
    const double dA = fabs(detJ) * mn * tb.weights[qa] * tb.weights[qb];
    const double *tq = a.traction + ((int64_t)f * N2 + t) * 3;
    for (int c = 0; c < 3; ++c) s_t[t * 3 + c] = tq[c] * nsn * dA;
  }
  __syncthreads();
  if (t < N2) {
    const int na = t % N, nb = t / N;
    int e3[3];
    e3[axis] = side ? N - 1 : 0;
    e3[ta] = na;
    e3[tbx] = nb;
    const uint32_t entry = map[e3[0] + N * (e3[1] + N * e3[2])];
    const int64_t node = entry & kNodeMaskDev;
    double acc[3] = {0, 0, 0};
    for (int qb = 0; qb < N; ++qb)
      for (int qa = 0; qa < N; ++qa) {
        const double phi = tb.val[qa * N + na] * tb.val[qb * N + nb];
        for (int c = 0; c < 3; ++c) acc[c] += s_t[(qa + N * qb) * 3 + c] * phi;
      }
    for (int c = 0; c < 3; ++c)
      if (!((entry >> (kConstraintShift + c)) & 1u)) atomicAdd(&a.rhs[node * 3 + c], acc[c]);
  }
}

// Values of an nc-component field at the face points (what the reference
// writes to preCICE at the interface quadrature points: the displacement of
// the solid, the temperature of the heat problem).  Only the face's nodes
// carry nonzero basis values there (GLL nodes).  One thread per face point.
__global__ void face_values_kernel(const uint32_t *__restrict__ cell_nodes, const int32_t *__restrict__ faces,
                                   int n_faces, int nc, FaceTables tb, const double *__restrict__ u,
                                   double *__restrict__ out) {
  const int N = tb.N, N2 = N * N, N3 = N2 * N;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)n_faces * N2) return;
  const int f = (int)(g / N2), q = (int)(g - (int64_t)f * N2), qa = q % N, qb = q / N;
  int axis, side, ta, tbx;
  face_axes(faces[2 * f + 1], axis, side, ta, tbx);
  const uint32_t *map = cell_nodes + (int64_t)faces[2 * f] * N3;
  double acc[3] = {0, 0, 0};
  for (int nb = 0; nb < N; ++nb)
    for (int na = 0; na < N; ++na) {
      int e3[3];
      e3[axis] = side ? N - 1 : 0;
      e3[ta] = na;
      e3[tbx] = nb;
      const int64_t node = map[e3[0] + N * (e3[1] + N * e3[2])] & kNodeMaskDev;
      const double phi = tb.val[qa * N + na] * tb.val[qb * N + nb];
      for (int c = 0; c < nc; ++c) acc[c] += phi * u[node * nc + c];
    }
  for (int c = 0; c < nc; ++c) out[g * nc + c] = acc[c];
}

// Heat interface flux (LaplaceProblem::assemble_rhs, heat_transfer.h:618-652):
// rhs -= dt int q phi dA over the faces.  One workgroup per face: the point
// weights -dt q dA go through LDS to one thread per face node; atomic sums on
// shared edge nodes; constrained dofs are skipped.
__global__ void __launch_bounds__(64) interface_flux_kernel(const uint32_t *__restrict__ cell_nodes,
                                                            const double *__restrict__ vertices,
                                                            const int32_t *__restrict__ faces, FaceTables tb,
                                                            const double *__restrict__ flux, double dt,
                                                            double *__restrict__ rhs) {
  const int N = tb.N, N2 = N * N, N3 = N2 * N;
  __shared__ double s_w[kFaceMaxN * kFaceMaxN];
  const int f = blockIdx.x, t = threadIdx.x;
  const int cell = faces[2 * f];
  int axis, side, ta, tbx;
  face_axes(faces[2 * f + 1], axis, side, ta, tbx);
  if (t < N2) {
    const int qa = t % N, qb = t / N;
    double xi[3];
    xi[axis] = side;
    xi[ta] = tb.points[qa];
    xi[tbx] = tb.points[qb];
    double J[3][3], Ji[3][3];
    trilinear_jacobian(vertices + (int64_t)cell * 24, xi[0], xi[1], xi[2], J);
    inv3(J, Ji);
    const double mn = sqrt(Ji[axis][0] * Ji[axis][0] + Ji[axis][1] * Ji[axis][1] + Ji[axis][2] * Ji[axis][2]);
    const double dA = fabs(det3(J)) * mn * tb.weights[qa] * tb.weights[qb];
    s_w[t] = -dt * flux[(int64_t)f * N2 + t] * dA;
  }
  __syncthreads();
  if (t < N2) {
    const int na = t % N, nb = t / N;
    int e3[3];
    e3[axis] = side ? N - 1 : 0;
    e3[ta] = na;
    e3[tbx] = nb;
    const uint32_t entry = cell_nodes[(int64_t)cell * N3 + e3[0] + N * (e3[1] + N * e3[2])];
    double acc = 0;
    for (int qb = 0; qb < N; ++qb)
      for (int qa = 0; qa < N; ++qa) acc += s_w[qa + N * qb] * tb.val[qa * N + na] * tb.val[qb * N + nb];
    if (!((entry >> kConstraintShift) & 1u)) atomicAdd(&rhs[entry & kNodeMaskDev], acc);
  }
}

// Reference coordinates of the face points, point q = qa + N qb.
__global__ void face_points_kernel(const double *__restrict__ vertices, const int32_t *__restrict__ faces, int n_faces,
                                   FaceTables tb, double *__restrict__ xyz) {
  const int N = tb.N, N2 = N * N;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)n_faces * N2) return;
  const int f = (int)(g / N2), q = (int)(g - (int64_t)f * N2);
  int axis, side, ta, tbx;
  face_axes(faces[2 * f + 1], axis, side, ta, tbx);
  double xi[3];
  xi[axis] = side;
  xi[ta] = tb.points[q % N];
  xi[tbx] = tb.points[q / N];
  const double *X = vertices + (int64_t)faces[2 * f] * 24;
  double x[3] = {0, 0, 0};
  for (int v = 0; v < 8; ++v) {
    double phi = 1;
    for (int d = 0; d < 3; ++d) phi *= ((v >> d) & 1) ? xi[d] : 1 - xi[d];
    for (int e = 0; e < 3; ++e) x[e] += phi * X[v * 3 + e];
  }
  for (int e = 0; e < 3; ++e) xyz[g * 3 + e] = x[e];
}

// Validates the (input cell, face) pairs and uploads them as (stored cell, face).
DeviceBuffer stored_faces(const OperatorData &op, int64_t n_faces, const int32_t *faces) {
  MFG_REQUIRE(n_faces >= 0 && n_faces < (int64_t)1 << 30, MFG_ERR_ARGUMENT, "bad face count");
  MFG_REQUIRE(n_faces == 0 || faces, MFG_ERR_ARGUMENT, "null face list");
  std::vector<int32_t> inv(op.n_cells);
  for (int c = 0; c < op.n_cells; ++c) inv[op.perm[c]] = c;
  std::vector<int32_t> h((size_t)n_faces * 2);
  for (int64_t f = 0; f < n_faces; ++f) {
    const int32_t cell = faces[2 * f], face = faces[2 * f + 1];
    MFG_REQUIRE(cell >= 0 && cell < op.n_cells && face >= 0 && face < 6, MFG_ERR_ARGUMENT, "bad (cell, face) pair");
    h[2 * f] = inv[cell];
    h[2 * f + 1] = face;
  }
  DeviceBuffer b;
  if (n_faces) {
    b.allocate(h.size() * sizeof(int32_t));
    MFG_HIP(hipMemcpyAsync(b.ptr, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice, op.stream));
    MFG_HIP(hipStreamSynchronize(op.stream));
  }
  return b;
}

void add_interface_traction(const OperatorData &op, int64_t n_faces, const int32_t *stored, const double *u,
                            const double *traction, double *rhs) {
  if (n_faces == 0) return;
  TractionArgs a{op.cell_nodes.as<uint32_t>(), op.vertices.as<double>(), stored, u, traction, rhs};
  interface_traction_kernel<<<(int)n_faces, 64, 0, op.stream>>>(a, face_tables(op.p));
  MFG_HIP(hipGetLastError());
}

}  // namespace mfg

using namespace mfg;

extern "C" {

int mfg_solid_add_interface_traction(mfg_operator op, int64_t n_faces, const int32_t *faces, const double *u,
                                     const double *traction, double *rhs) {
  return guarded([&] {
    MFG_REQUIRE(op, MFG_ERR_ARGUMENT, "null operator");
    MFG_REQUIRE(op->kind == 0 && op->precision == MFG_FP64, MFG_ERR_ARGUMENT,
                "the interface traction is assembled by FP64 solid operators");
    MFG_REQUIRE(u && rhs && (n_faces == 0 || traction), MFG_ERR_ARGUMENT, "null vector");
    DeviceBuffer fb = stored_faces(*op, n_faces, faces);
    if (n_faces == 0) return;
    add_interface_traction(*op, n_faces, fb.as<int32_t>(), u, traction, rhs);
    MFG_HIP(hipStreamSynchronize(op->stream));  // the face list is freed on return
  });
}

int mfg_operator_face_quadrature_points(mfg_operator op, int64_t n_faces, const int32_t *faces, double *xyz) {
  return guarded([&] {
    MFG_REQUIRE(op && (xyz || n_faces == 0), MFG_ERR_ARGUMENT, "null argument");
    DeviceBuffer fb = stored_faces(*op, n_faces, faces);
    if (n_faces == 0) return;
    const FaceTables tb = face_tables(op->p);
    const int64_t n = n_faces * tb.N * tb.N;
    face_points_kernel<<<blocks_for(n, 256), 256, 0, op->stream>>>(op->vertices.as<double>(), fb.as<int32_t>(),
                                                                    (int)n_faces, tb, xyz);
    MFG_HIP(hipGetLastError());
    MFG_HIP(hipStreamSynchronize(op->stream));
  });
}

int mfg_operator_face_values(mfg_operator op, int64_t n_faces, const int32_t *faces, const double *u,
                             double *values) {
  return guarded([&] {
    MFG_REQUIRE(op && u && (values || n_faces == 0), MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(op->nc <= 3, MFG_ERR_ARGUMENT, "at most 3 components");
    DeviceBuffer fb = stored_faces(*op, n_faces, faces);
    if (n_faces == 0) return;
    const FaceTables tb = face_tables(op->p);
    const int64_t n = n_faces * tb.N * tb.N;
    face_values_kernel<<<blocks_for(n, 256), 256, 0, op->stream>>>(op->cell_nodes.as<uint32_t>(), fb.as<int32_t>(),
                                                                    (int)n_faces, op->nc, tb, u, values);
    MFG_HIP(hipGetLastError());
    MFG_HIP(hipStreamSynchronize(op->stream));
  });
}

int mfg_heat_add_interface_flux(mfg_operator op, int64_t n_faces, const int32_t *faces, const double *flux,
                                double *rhs) {
  return guarded([&] {
    MFG_REQUIRE(op && rhs && (flux || n_faces == 0), MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(op->kind == 1, MFG_ERR_ARGUMENT, "the interface flux belongs to the heat operator");
    DeviceBuffer fb = stored_faces(*op, n_faces, faces);
    if (n_faces == 0) return;
    interface_flux_kernel<<<(int)n_faces, 64, 0, op->stream>>>(op->cell_nodes.as<uint32_t>(), op->vertices.as<double>(),
                                                               fb.as<int32_t>(), face_tables(op->p), flux, op->dt, rhs);
    MFG_HIP(hipGetLastError());
    MFG_HIP(hipStreamSynchronize(op->stream));
  });
}

}  // extern "C"
