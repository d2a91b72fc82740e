// Geometric multigrid transfer between nested hex levels (MGTransferMatrixFree
// as used at mf_elasticity.h:863-877, 920-922 and heat_transfer.h:761-762):
// prolongation is the embedding of the coarse FE space (values of the coarse
// function at the fine support points), restriction its exact transpose, and
// interpolation the nodal interpolant of a fine function on the coarse level.
//
// One workgroup per coarse cell.  The (2p+1)^3 fine support points inside a
// coarse cell form a lattice; the 1D embedding matrix E[(2p+1) x (p+1)] is
// applied in three LDS sweeps.  Every fine lattice entry carries the fine node
// index, the fine constraint bits (28..30) and an owner bit (31): exactly one
// coarse cell owns each fine node, so prolongation adds each fine value once
// without atomics, and restriction counts each fine value once.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cell_kernels.hpp"

namespace mfg {

constexpr uint32_t kOwnerBit = 1u << 31;
constexpr uint32_t kFineNodeMask = (1u << 28) - 1u;

template <typename CN, typename FN>
struct TransferArgs {
  const uint32_t *coarse_nodes;  // coarse map (stored cell order), constraint bits
  const uint32_t *lattice;       // (2p+1)^3 per coarse cell, same cell order
  int cell_begin, cell_end;
  const double *E;               // (2p+1) x (p+1), row-major
  const CN *coarse_in;
  CN *coarse_out;
  const FN *fine_in;
  FN *fine_out;
};

// LDS doubles one coarse cell needs (the largest of the three transfers)
template <int P>
constexpr int transfer_lds() {
  constexpr int N = P + 1, M = 2 * P + 1;
  return M * N + M * M * M + M * M * N + M * N * N;
}

// The per-cell bodies run on all threads of a workgroup (strided loops; the
// sizes are small and vary with p) and use `lds` as scratch.
template <int P, int NC, typename CN, typename FN>
__device__ inline void prolongate_cell(const TransferArgs<CN, FN> &a, int cell, double *lds) {
  constexpr int N = P + 1, M = 2 * P + 1, N3 = N * N * N, M3 = M * M * M;
  double *sE = lds, *su = sE + M * N, *t1 = su + N3, *t2 = t1 + M * N * N;
  const int t = threadIdx.x, nt = blockDim.x;
  for (int x = t; x < M * N; x += nt) sE[x] = a.E[x];
  for (int c = 0; c < NC; ++c) {
    __syncthreads();
    for (int x = t; x < N3; x += nt) {
      const uint32_t e = a.coarse_nodes[(int64_t)cell * N3 + x];
      const bool con = (e >> (kConstraintShift + c)) & 1u;
      su[x] = con ? 0.0 : (double)a.coarse_in[(int64_t)(e & kNodeMaskDev) * NC + c];
    }
    __syncthreads();
    for (int x = t; x < M * N * N; x += nt) {  // t1[I][b][c'] (I fastest)
      const int I = x % M, bc = x / M;
      double s = 0;
      for (int q = 0; q < N; ++q) s += sE[I * N + q] * su[q + N * bc];
      t1[x] = s;
    }
    __syncthreads();
    for (int x = t; x < M * M * N; x += nt) {  // t2[I][J][c']
      const int I = x % M, J = (x / M) % M, cz = x / (M * M);
      double s = 0;
      for (int q = 0; q < N; ++q) s += sE[J * N + q] * t1[I + M * (q + N * cz)];
      t2[x] = s;
    }
    __syncthreads();
    for (int x = t; x < M3; x += nt) {
      const uint32_t e = a.lattice[(int64_t)cell * M3 + x];
      if (!(e & kOwnerBit) || ((e >> (kConstraintShift + c)) & 1u)) continue;
      const int IJ = x % (M * M), K = x / (M * M);
      double s = 0;
      for (int q = 0; q < N; ++q) s += sE[K * N + q] * t2[IJ + M * M * q];
      FN *dst = a.fine_out + (int64_t)(e & kFineNodeMask) * NC + c;
      *dst = (FN)((double)*dst + s);
    }
  }
}

// coarse += P^T fine on one coarse cell (callers run one colour at a time)
template <int P, int NC, typename CN, typename FN>
__device__ inline void restrict_cell(const TransferArgs<CN, FN> &a, int cell, double *lds) {
  constexpr int N = P + 1, M = 2 * P + 1, N3 = N * N * N, M3 = M * M * M;
  double *sE = lds, *sf = sE + M * N, *t2 = sf + M3, *t1 = t2 + M * M * N;
  const int t = threadIdx.x, nt = blockDim.x;
  for (int x = t; x < M * N; x += nt) sE[x] = a.E[x];
  for (int c = 0; c < NC; ++c) {
    __syncthreads();
    for (int x = t; x < M3; x += nt) {
      const uint32_t e = a.lattice[(int64_t)cell * M3 + x];
      const bool use = (e & kOwnerBit) && !((e >> (kConstraintShift + c)) & 1u);
      sf[x] = use ? (double)a.fine_in[(int64_t)(e & kFineNodeMask) * NC + c] : 0.0;
    }
    __syncthreads();
    for (int x = t; x < M * M * N; x += nt) {  // t2[I][J][k]
      const int IJ = x % (M * M), k = x / (M * M);
      double s = 0;
      for (int q = 0; q < M; ++q) s += sE[q * N + k] * sf[IJ + M * M * q];
      t2[x] = s;
    }
    __syncthreads();
    for (int x = t; x < M * N * N; x += nt) {  // t1[I][j][k]
      const int I = x % M, j = (x / M) % N, k = x / (M * N);
      double s = 0;
      for (int q = 0; q < M; ++q) s += sE[q * N + j] * t2[I + M * (q + M * k)];
      t1[x] = s;
    }
    __syncthreads();
    for (int x = t; x < N3; x += nt) {
      const uint32_t e = a.coarse_nodes[(int64_t)cell * N3 + x];
      if ((e >> (kConstraintShift + c)) & 1u) continue;
      const int i = x % N, jk = x / N;
      double s = 0;
      for (int q = 0; q < M; ++q) s += sE[q * N + i] * t1[q + M * jk];
      CN *dst = a.coarse_out + (int64_t)(e & kNodeMaskDev) * NC + c;
      *dst = (CN)((double)*dst + s);
    }
  }
}

// coarse = nodal interpolant of the fine function; E here is the
// interpolation matrix (p+1) x (2p+1).  Shared coarse nodes receive
// bit-identical values from every cell containing them.
template <int P, int NC, typename CN, typename FN>
__device__ inline void interpolate_cell(const TransferArgs<CN, FN> &a, int cell, double *lds) {
  constexpr int N = P + 1, M = 2 * P + 1, N3 = N * N * N, M3 = M * M * M;
  double *sI = lds, *sf = sI + N * M, *t2 = sf + M3, *t1 = t2 + M * M * N;
  const int t = threadIdx.x, nt = blockDim.x;
  for (int x = t; x < N * M; x += nt) sI[x] = a.E[x];
  for (int c = 0; c < NC; ++c) {
    __syncthreads();
    for (int x = t; x < M3; x += nt) {
      const uint32_t e = a.lattice[(int64_t)cell * M3 + x];
      sf[x] = (double)a.fine_in[(int64_t)(e & kFineNodeMask) * NC + c];
    }
    __syncthreads();
    for (int x = t; x < M * M * N; x += nt) {  // t2[I][J][k]
      const int IJ = x % (M * M), k = x / (M * M);
      double s = 0;
      for (int q = 0; q < M; ++q) s += sI[k * M + q] * sf[IJ + M * M * q];
      t2[x] = s;
    }
    __syncthreads();
    for (int x = t; x < M * N * N; x += nt) {
      const int I = x % M, j = (x / M) % N, k = x / (M * N);
      double s = 0;
      for (int q = 0; q < M; ++q) s += sI[j * M + q] * t2[I + M * (q + M * k)];
      t1[x] = s;
    }
    __syncthreads();
    for (int x = t; x < N3; x += nt) {
      const uint32_t e = a.coarse_nodes[(int64_t)cell * N3 + x];
      const int i = x % N, jk = x / N;
      double s = 0;
      for (int q = 0; q < M; ++q) s += sI[i * M + q] * t1[q + M * jk];
      a.coarse_out[(int64_t)(e & kNodeMaskDev) * NC + c] = (CN)s;
    }
  }
}

// One workgroup of 256 threads per coarse cell of [cell_begin, cell_end).
template <int P, int NC, typename CN, typename FN>
__global__ void __launch_bounds__(256) prolongate_add_kernel(TransferArgs<CN, FN> a) {
  __shared__ double lds[transfer_lds<P>()];
  const int cell = a.cell_begin + blockIdx.x;
  if (cell < a.cell_end) prolongate_cell<P, NC, CN, FN>(a, cell, lds);
}
template <int P, int NC, typename CN, typename FN>
__global__ void __launch_bounds__(256) restrict_add_kernel(TransferArgs<CN, FN> a) {
  __shared__ double lds[transfer_lds<P>()];
  const int cell = a.cell_begin + blockIdx.x;
  if (cell < a.cell_end) restrict_cell<P, NC, CN, FN>(a, cell, lds);
}
template <int P, int NC, typename CN, typename FN>
__global__ void __launch_bounds__(256) interpolate_kernel(TransferArgs<CN, FN> a) {
  __shared__ double lds[transfer_lds<P>()];
  const int cell = a.cell_begin + blockIdx.x;
  if (cell < a.cell_end) interpolate_cell<P, NC, CN, FN>(a, cell, lds);
}

}  // namespace mfg
