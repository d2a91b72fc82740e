// Shared host/device plumbing of libmfgpu: status handling, device buffers,
// the 1D finite-element tables and launch helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mfgpu.h"

namespace mfg {

// Error carried to the C ABI as a status code plus a thread-local message.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
void set_last_error(const std::string &msg);

#define MFG_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      throw ::mfg::Error(MFG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define MFG_REQUIRE(cond, code, msg)                                                           \
  do {                                                                                         \
    if (!(cond)) throw ::mfg::Error((code), (msg));                                            \
  } while (0)

// Runs f, mapping exceptions to a status code (the ABI never throws).
template <class F>
int guarded(F &&f) {
  try {
    f();
    return MFG_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return MFG_ERR_ARGUMENT;
  }
}

// Owning device allocation.
struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t b) { allocate(b); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  DeviceBuffer(DeviceBuffer &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
  DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
    if (this != &o) { release(); ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DeviceBuffer() { release(); }
  void allocate(size_t b) {
    release();
    if (b == 0) return;
    MFG_HIP(hipMalloc(&ptr, b));
    bytes = b;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T> T *as() const { return static_cast<T *>(ptr); }
};

template <class T>
inline void upload(DeviceBuffer &buf, const T *host, size_t n, hipStream_t s) {
  buf.allocate(n * sizeof(T));
  if (n) {
    MFG_HIP(hipMemcpyAsync(buf.ptr, host, n * sizeof(T), hipMemcpyHostToDevice, s));
    MFG_HIP(hipStreamSynchronize(s));
  }
}

// ------------------------------------------------ 1D finite element tables ---
// Support points of FE_Q(p): Gauss-Lobatto points on [0,1].
std::vector<double> gauss_lobatto_points(int p);
// QGauss(n) on [0,1].
void gauss_points(int n, std::vector<double> &x, std::vector<double> &w);
// Lagrange basis of nodes z at point x (barycentric form): values, derivatives.
void lagrange_basis(const std::vector<double> &z, double x, double *val, double *der);

// Tables the sum-factorisation kernels need for degree p with n_q = p+1:
//   shape[q][a]  = phi_a(xi_q)         (GLL basis at the Gauss points)
//   colloc[q][r] = l_r'(xi_q)          (derivative of the Gauss-point Lagrange
//                                       basis: the collocation derivative)
//   weight[q]    = Gauss weight
struct Tables1D {
  int N = 0;
  std::vector<double> shape, colloc, weight, points;
};
Tables1D make_tables(int p);

// Maximum supported degree of the device kernels.
constexpr int kMaxDegree = 6;
// Node indices are stored with constraint bits for the components above bit 28.
constexpr int kNodeBits = 28;
constexpr uint32_t kNodeMask = (1u << kNodeBits) - 1u;

inline int blocks_for(int64_t n, int threads) { return (int)((n + threads - 1) / threads); }

}  // namespace mfg
