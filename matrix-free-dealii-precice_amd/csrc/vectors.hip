// Synthetic input generation on the device (seeded, reproducible).
#include "common.hpp"
#include "vector_kernels.hpp"

namespace mfg {

__host__ __device__ inline uint64_t splitmix64(uint64_t seed, uint64_t index) {
  uint64_t z = seed * 0xD1B54A32D192ED03ull + (index + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
__global__ void uniform_kernel(uint64_t seed, int64_t n, T *dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double u = (double)(splitmix64(seed, (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0);
    dst[i] = (T)(2.0 * u - 1.0);
  }
}

double uniform_pm1_host(uint64_t seed, uint64_t index) {
  return 2.0 * (double)(splitmix64(seed, index) >> 11) * (1.0 / 9007199254740992.0) - 1.0;
}

}  // namespace mfg

using namespace mfg;

extern "C" int mfg_fill_uniform(uint64_t seed, int64_t n, int32_t precision, void *dst, void *stream) {
  return guarded([&] {
    MFG_REQUIRE(n >= 0 && (dst || n == 0), MFG_ERR_ARGUMENT, "bad vector");
    if (n == 0) return;
    const int threads = 256;
    const int blocks = (int)std::min<int64_t>((n + threads - 1) / threads, 65536);
    if (precision == MFG_FP64)
      uniform_kernel<double><<<blocks, threads, 0, (hipStream_t)stream>>>(seed, n, (double *)dst);
    else
      uniform_kernel<float><<<blocks, threads, 0, (hipStream_t)stream>>>(seed, n, (float *)dst);
    MFG_HIP(hipGetLastError());
  });
}
