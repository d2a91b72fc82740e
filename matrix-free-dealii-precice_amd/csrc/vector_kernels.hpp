// Element-wise and reduction kernels on level vectors.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mfg {

template <typename T>
__global__ void fill_kernel(T *x, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

template <typename T>
__global__ void set_entries_kernel(T *x, const int32_t *idx, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[idx[i]] = v;
}

// inv = 1/d; flags non-positive entries
template <typename T>
__global__ void invert_kernel(const T *d, T *inv, int64_t n, int *flag) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T v = d[i];
    if (!(v > T(0))) atomicOr(flag, 1);
    inv[i] = T(1) / v;
  }
}

// Per-(cell, point) field arrays from colour-major to input cell order:
// out[f][perm[c]*n3 + q] = in[f][c*n3 + q]
template <typename T>
__global__ void unpermute_fields_kernel(const T *in, T *out, const int32_t *perm, int64_t n_cells, int n3,
                                        int n_fields) {
  const int64_t nqp = n_cells * n3;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nqp * n_fields;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = i / nqp, r = i - f * nqp, c = r / n3, q = r - c * n3;
    out[f * nqp + (int64_t)perm[c] * n3 + q] = in[i];
  }
}

// dst = omega * a .* b
template <typename T>
__global__ void scaled_product_kernel(T *dst, const T *a, const T *b, T omega, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = omega * a[i] * b[i];
}

}  // namespace mfg
