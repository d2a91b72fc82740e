// Element-wise and reduction kernels on level vectors.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include "smoother_terms.hpp"

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
// out[f][perm[c]*n3 + q] = in[f][c*n3 + q]; paired: the input holds the
// fields in pairs, in[((f/2) nqp + r) 2 + f%2], an odd last field plainly
// after them
template <typename T>
__global__ void unpermute_fields_kernel(const T *in, T *out, const int32_t *perm, int64_t n_cells, int n3,
                                        int n_fields, int paired) {
  const int64_t nqp = n_cells * n3;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nqp * n_fields;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = i / nqp, r = i - f * nqp, c = r / n3, q = r - c * n3;
    const bool in_pair = paired && f < (n_fields & ~1);
    out[f * nqp + (int64_t)perm[c] * n3 + q] = in[in_pair ? ((f >> 1) * nqp + r) * 2 + (f & 1) : i];
  }
}

#define MFG_GRID_STRIDE(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

template <typename TO, typename TI>
__global__ void convert_kernel(TO *dst, const TI *src, int64_t n) {
  MFG_GRID_STRIDE(i, n) dst[i] = (TO)src[i];
}

// Deterministic two-pass reduction: fixed grid, per-block tree in LDS, then a
// single block sums the partials in a fixed order.
constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

template <typename T>
__device__ inline void block_reduce_store(double v, double *partials) {
  __shared__ double s[kRedThreads];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads) dot_partial_kernel(const T *a, const T *b, int64_t n, double *partials) {
  double v = 0;
  MFG_GRID_STRIDE(i, n) v += (double)a[i] * (double)b[i];
  block_reduce_store<T>(v, partials);
}

static __global__ void __launch_bounds__(kRedThreads) sum_partials_kernel(const double *partials, int n, double *out) {
  __shared__ double s[kRedThreads];
  double v = 0;
  for (int i = threadIdx.x; i < n; i += kRedThreads) v += partials[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// x += alpha p ; r -= alpha h ; partial r.r
template <typename T>
__global__ void __launch_bounds__(kRedThreads) cg_update_kernel(T *x, T *r, const T *p, const T *h, double alpha,
                                                                 int64_t n, double *partials) {
  double v = 0;
  MFG_GRID_STRIDE(i, n) {
    x[i] = (T)((double)x[i] + alpha * (double)p[i]);
    const T ri = (T)((double)r[i] - alpha * (double)h[i]);
    r[i] = ri;
    v += (double)ri * (double)ri;
  }
  block_reduce_store<T>(v, partials);
}

// p = d + beta p
template <typename T>
__global__ void xpby_kernel(T *p, const T *d, double beta, int64_t n) {
  MFG_GRID_STRIDE(i, n) p[i] = (T)((double)d[i] + beta * (double)p[i]);
}

// t = b - t
template <typename T>
__global__ void residual_kernel(T *t, const T *b, int64_t n) {
  MFG_GRID_STRIDE(i, n) t[i] = b[i] - t[i];
}

// y = a + s*b   (general axpby into a third vector)
template <typename T>
__global__ void add_scaled_kernel(T *y, const T *a, double s, const T *b, int64_t n) {
  MFG_GRID_STRIDE(i, n) y[i] = (T)((double)a[i] + s * (double)b[i]);
}

// Chebyshev (zero initial guess): d = D^{-1} b / theta ; x = d
template <typename T>
__global__ void cheb_first_kernel(T *x, T *d, const T *b, const T *invD, double inv_theta, int64_t n) {
  MFG_GRID_STRIDE(i, n) {
    const T v = (T)((double)invD[i] * (double)b[i] * inv_theta);
    d[i] = v;
    x[i] = v;
  }
}
// Chebyshev (given x, t = A x): d = D^{-1}(b - t)/theta ; x += d
template <typename T>
__global__ void cheb_first_step_kernel(T *x, T *d, const T *b, const T *t, const T *invD, double inv_theta,
                                       int64_t n) {
  MFG_GRID_STRIDE(i, n) {
    const T v = cheb_first_step_value<T>(invD[i], b[i], t[i], inv_theta);
    d[i] = v;
    x[i] += v;
  }
}
// Chebyshev recurrence: d = f1 d + f2 D^{-1}(b - t) ; x += d
template <typename T>
__global__ void cheb_update_kernel(T *x, T *d, const T *b, const T *t, const T *invD, double f1, double f2,
                                   int64_t n) {
  MFG_GRID_STRIDE(i, n) {
    const T v = cheb_update_value<T>(d[i], invD[i], b[i], t[i], f1, f2);
    d[i] = v;
    x[i] += v;
  }
}

// dst = omega * a .* b
template <typename T>
__global__ void scaled_product_kernel(T *dst, const T *a, const T *b, T omega, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = omega * a[i] * b[i];
}

}  // namespace mfg
