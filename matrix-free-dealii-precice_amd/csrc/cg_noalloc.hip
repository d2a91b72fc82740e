// Allocation-free CG (mfg_cg_solve_noalloc): the entry point of
// mfg_cg_solve for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops).  Its reduction partials, device
// scalars and CG vectors live in a file-scope device array; nothing here calls
// hipMalloc or hipFree.
#include <cstring>
#include <mutex>

#include "common.hpp"
#include "operator.hpp"
#include "solver.hpp"
#include "vector_kernels.hpp"

using namespace mfg;

// ------------------------------------------- allocation-free CG scratch ---
// Scratch of mfg_cg_solve_noalloc: reduction partials, the scalar block, then
// the four CG vectors r, z, p, h of up to kNoAllocMaxDofs entries (FP64), each
// starting on a 256-byte boundary.  The array is module data: every GPU the
// library runs on holds its own copy (32 MiB + 8 KiB) until the process ends.
static constexpr int64_t kNoAllocMaxDofs = (int64_t)1 << 20;
static constexpr size_t kNoAllocHead = (kRedBlocks + kRedScalars) * sizeof(double);
static_assert(kNoAllocHead % 256 == 0, "vectors start on 256-byte boundaries");
static constexpr size_t kNoAllocBytes = kNoAllocHead + 4 * (size_t)kNoAllocMaxDofs * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// The host-side name `scratch` is not a device address: its device address is
// looked up once per device, with that device current, and kept.
static constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_dev[kMaxScratchDevices] = {};

static unsigned char *scratch_for_current_device() {
  int dev = -1;
  MFG_HIP(hipGetDevice(&dev));
  MFG_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, MFG_ERR_ARGUMENT, "device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    MFG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    MFG_REQUIRE(ptr != nullptr, MFG_ERR_HIP, "no device address for the CG scratch");
    g_scratch_dev[dev] = (unsigned char *)ptr;
  }
  return g_scratch_dev[dev];
}

extern "C" {

int mfg_cg_solve_noalloc(mfg_operator A, int32_t preconditioner, mfg_mg mg, double omega, const void *b, void *x,
                         double tol_rel, int32_t max_iter, mfg_solve_info *info) {
  mfg_solve_info dummy;
  if (!info) info = &dummy;
  std::memset(info, 0, sizeof(*info));
  return guarded([&] {
    MFG_REQUIRE(A && b && x, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(b != x, MFG_ERR_ARGUMENT, "b and x must be different vectors");
    MFG_REQUIRE(preconditioner >= 0 && preconditioner <= 2, MFG_ERR_ARGUMENT, "unknown preconditioner");
    MFG_REQUIRE(A->n_dofs <= kNoAllocMaxDofs, MFG_ERR_ARGUMENT,
                "mfg_cg_solve_noalloc: the operator has more than 2^20 rows (the scratch limit); use mfg_cg_solve");
    // a graphed V-cycle re-captures (and allocates) after every set-up
    MFG_REQUIRE(preconditioner != MFG_PRECOND_MG || !mg || !mg->use_graph, MFG_ERR_ARGUMENT,
                "mfg_cg_solve_noalloc: the multigrid preconditioner must not replay graphs");
    unsigned char *base = scratch_for_current_device();
    Reducer red((double *)base, (double *)base + kRedBlocks);
    const size_t vb = ((size_t)A->n_dofs * A->elem_size() + 255) & ~(size_t)255;
    unsigned char *v0 = base + kNoAllocHead;
    cg_solve(*A, preconditioner, mg, omega, b, x, tol_rel, max_iter, info,
             CGVectors{v0, v0 + vb, v0 + 2 * vb, v0 + 3 * vb}, red);
  });
}

}  // extern "C"
