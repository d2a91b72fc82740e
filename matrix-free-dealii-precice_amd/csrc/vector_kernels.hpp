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

// The reduction kernels take kRedBatch grid-stride iterations at a time while
// a whole batch is in range: the loads of the batch issue before the first
// use (more bytes in flight per wave); the rest go one by one.  Each thread
// still visits its indices, and sums, in increasing order, so the partial
// sums are those of the plain grid-stride loop.  load(i, u) fills slot u,
// use(i, u) consumes it.
constexpr int kRedBatch = 4;
template <typename Load, typename Use>
__device__ inline void red_batches(int64_t n, Load load, Use use) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (kRedBatch - 1) * stride < n; i += kRedBatch * stride) {
#pragma unroll
    for (int u = 0; u < kRedBatch; ++u) load(i + u * stride, u);
#pragma unroll
    for (int u = 0; u < kRedBatch; ++u) use(i + u * stride, u);
  }
  for (; i < n; i += stride) {
    load(i, 0);
    use(i, 0);
  }
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads) dot_partial_kernel(const T *a, const T *b, int64_t n, double *partials) {
  double v = 0;
  T av[kRedBatch], bv[kRedBatch];
  red_batches(
      n,
      [&](int64_t i, int u) {
        av[u] = a[i];
        bv[u] = b[i];
      },
      [&](int64_t, int u) { v += (double)av[u] * (double)bv[u]; });
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

// ------------------------------------------ device-resident CG scalars ---
// A block of scalars that the reduction epilogues and the update kernels read
// and write, so that a CG iteration computes its step lengths on the device
// and the host copies back only what it decides on (the residual norm, once
// per iteration) or nothing at all (fixed-count runs, stopped by a flag).
enum RedSlot : int {
  kSlotGH = 0,     // (r, z) of the current iterate
  kSlotPH,         // (p, A p)
  kSlotAlpha,      // step length of this iteration; 0 when the iteration is inactive
  kSlotBeta,       // direction-update coefficient of the last active iteration
  kSlotRR,         // (r, r) after the update
  kSlotB2,         // (b, b)
  kSlotRes0,       // ||r_0|| of a fixed-count run
  kSlotTol,        // a fixed-count run stops at ||r|| <= max(tol, rel ||r_0||)
  kSlotRel,
  kSlotStop,       // 1: later iterations change neither x, r nor the history
  kSlotActive,     // 1: this iteration took a step
  kSlotBreakdown,  // 1: (p, A p) <= 0 was met
  kSlotNAlpha,     // history lengths (integers stored as doubles)
  kSlotNBeta,
  kSlotHost,       // host-synchronous dot products
  kSlotUpd,        // (du, du) of the Newton update, read with the next residual
  kSlotNewton,     // (R, R) of the Newton residual
  kRedScalars = 32
};
// Epilogue of the final reduction pass (v = the reduced sum):
//   kRedSet     scal[slot] = v
//   kRedStart   (r, z) of the first iterate: resets the flags and the history
//   kRedRes0    ||r_0|| = sqrt(v) and the stop threshold (tol, rel); stop at once if ||r_0|| <= max(tol, rel ||r_0||)
//   kRedAlpha   (p, A p): alpha = gh / v and record it, or stop on breakdown
//   kRedRRStop  (r, r) after the update; stop when below the threshold
//   kRedBeta    (r, z) of the new iterate: beta = v / gh and record it
enum RedMode : int { kRedSet = 0, kRedStart, kRedRes0, kRedAlpha, kRedRRStop, kRedBeta };

__device__ inline void red_epilogue(double v, double *scal, int mode, int slot, double tol, double rel,
                                    double *hist, int hist_cap) {
  switch (mode) {
    case kRedSet: scal[slot] = v; break;
    case kRedStart:
      scal[kSlotGH] = v;
      scal[kSlotStop] = 0;
      scal[kSlotActive] = 1;
      scal[kSlotBreakdown] = 0;
      scal[kSlotNAlpha] = 0;
      scal[kSlotNBeta] = 0;
      scal[kSlotTol] = 0;
      scal[kSlotRel] = 0;
      scal[kSlotAlpha] = 0;
      scal[kSlotBeta] = 0;
      break;
    case kRedRes0: {
      const double r0 = sqrt(v);
      scal[kSlotRes0] = r0;
      scal[kSlotTol] = tol;
      scal[kSlotRel] = rel;
      if (!(r0 > fmax(tol, rel * r0))) scal[kSlotStop] = 1;
      break;
    }
    case kRedAlpha:
      scal[kSlotPH] = v;
      if (scal[kSlotStop] != 0) {
        scal[kSlotAlpha] = 0;
        scal[kSlotActive] = 0;
      } else if (!(v > 0)) {
        scal[kSlotAlpha] = 0;
        scal[kSlotActive] = 0;
        scal[kSlotStop] = 1;
        scal[kSlotBreakdown] = 1;
      } else {
        const double a = scal[kSlotGH] / v;
        scal[kSlotAlpha] = a;
        scal[kSlotActive] = 1;
        const int k = (int)scal[kSlotNAlpha];
        if (hist && k < hist_cap) {
          hist[k] = a;
          scal[kSlotNAlpha] = k + 1;
        }
      }
      break;
    case kRedRRStop:
      scal[kSlotRR] = v;
      if (scal[kSlotActive] != 0 && sqrt(v) <= fmax(scal[kSlotTol], scal[kSlotRel] * scal[kSlotRes0]))
        scal[kSlotStop] = 1;
      break;
    case kRedBeta:
      if (scal[kSlotActive] != 0 && scal[kSlotStop] == 0) {
        const double b = v / scal[kSlotGH];
        scal[kSlotGH] = v;
        scal[kSlotBeta] = b;
        const int k = (int)scal[kSlotNBeta];
        if (hist && k < hist_cap) {
          hist[hist_cap + k] = b;
          scal[kSlotNBeta] = k + 1;
        }
      } else {
        scal[kSlotBeta] = 0;
      }
      break;
  }
}

static __global__ void __launch_bounds__(kRedThreads)
    sum_partials_epilogue_kernel(const double *partials, int n, double *scal, int mode, int slot, double tol,
                                 double rel, double *hist, int hist_cap) {
  __shared__ double s[kRedThreads];
  double v = 0;
  for (int i = threadIdx.x; i < n; i += kRedThreads) v += partials[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) red_epilogue(s[0], scal, mode, slot, tol, rel, hist, hist_cap);
}

// x += alpha p ; r -= alpha h with alpha = scal[kSlotAlpha] (no change when it
// is 0, i.e. the iteration is inactive) ; partial r.r
template <typename T>
__global__ void __launch_bounds__(kRedThreads) cg_update_dev_kernel(T *x, T *r, const T *p, const T *h,
                                                                     const double *scal, int64_t n,
                                                                     double *partials) {
  const double alpha = scal[kSlotAlpha];
  double v = 0;
  T rv[kRedBatch], xv[kRedBatch], pv[kRedBatch], hv[kRedBatch];
  if (alpha != 0)
    red_batches(
        n,
        [&](int64_t i, int u) {
          rv[u] = r[i];
          xv[u] = x[i];
          pv[u] = p[i];
          hv[u] = h[i];
        },
        [&](int64_t i, int u) {
          x[i] = (T)((double)xv[u] + alpha * (double)pv[u]);
          const T ri = (T)((double)rv[u] - alpha * (double)hv[u]);
          r[i] = ri;
          v += (double)ri * (double)ri;
        });
  else
    red_batches(
        n, [&](int64_t i, int u) { rv[u] = r[i]; },
        [&](int64_t, int u) { v += (double)rv[u] * (double)rv[u]; });
  block_reduce_store<T>(v, partials);
}

// z = omega D^{-1} r ; partial r.z: the Jacobi step and the (r, z) product of
// a preconditioned CG in one pass, with the values of scaled_product_kernel
// and the partial sums of dot_partial_kernel (same expressions, same grid)
template <typename T>
__global__ void __launch_bounds__(kRedThreads) jacobi_dot_kernel(T *z, const T *invD, const T *r, T omega, int64_t n,
                                                                  double *partials) {
  double v = 0;
  T rv[kRedBatch], dv[kRedBatch];
  red_batches(
      n,
      [&](int64_t i, int u) {
        rv[u] = r[i];
        dv[u] = invD[i];
      },
      [&](int64_t i, int u) {
        const T zi = omega * dv[u] * rv[u];
        z[i] = zi;
        v += (double)rv[u] * (double)zi;
      });
  block_reduce_store<T>(v, partials);
}

// p = d + beta p with beta = scal[kSlotBeta]
template <typename T>
__global__ void xpby_dev_kernel(T *p, const T *d, const double *scal, int64_t n) {
  const double beta = scal[kSlotBeta];
  MFG_GRID_STRIDE(i, n) p[i] = (T)((double)d[i] + beta * (double)p[i]);
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
