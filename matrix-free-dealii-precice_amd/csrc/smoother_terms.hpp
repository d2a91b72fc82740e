// The Chebyshev smoother's per-entry update terms (PreconditionChebyshev's
// recurrence, configured at mf_elasticity.h:1149-1180).  The single-level
// vector kernels and the batched segment kernels both call these functions,
// with floating-point contraction off inside, so batched and per-instance
// V-cycles produce the same bits whatever each kernel's code generation.
#pragma once
#include <hip/hip_runtime.h>

namespace mfg {

// first step from a given x: D^{-1} (b - t) / theta
template <typename T>
__device__ inline T cheb_first_step_value(T inv_diag, T b, T t, double inv_theta) {
#pragma clang fp contract(off)
  return (T)((double)inv_diag * ((double)b - (double)t) * inv_theta);
}

// recurrence: f1 d + f2 D^{-1} (b - t)
template <typename T>
__device__ inline T cheb_update_value(T d, T inv_diag, T b, T t, double f1, double f2) {
#pragma clang fp contract(off)
  return (T)(f1 * (double)d + f2 * (double)inv_diag * ((double)b - (double)t));
}

}  // namespace mfg
