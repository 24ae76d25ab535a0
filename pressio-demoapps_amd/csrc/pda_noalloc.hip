// pda_noalloc.hip -- applyJacobian for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  pda_apply_jacobian assembles J(U) into a per-problem
// value buffer that it allocates on first use; this variant assembles it into
// one fixed file-scope device array instead and then runs the same kernels.
#include <mutex>

#include "pda_internal.hpp"

namespace pda {
namespace {

// Jacobian values of one call: at most PDA_NOALLOC_MAX_NNZ doubles (64 MiB),
// held by every device the library runs on until the process ends.
constexpr int64_t kScratchMax = int64_t(1) << 23;
__device__ __align__(256) double scratch[kScratchMax];

// Device address of `scratch` per device.  In host code the array's name is
// a host-side stand-in, never a device address: kernels and copies get the
// pointer from hipGetSymbolAddress, looked up once per device with that
// device current.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};

int scratch_on_current_device(double **out) {
  int dev = -1;
  PDA_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("pda_apply_jacobian_noalloc: device ordinal out of range");
    return PDA_E_HIP;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    PDA_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) {
      set_error("pda_apply_jacobian_noalloc: no device address for the scratch array");
      return PDA_E_HIP;
    }
    g_scratch[dev] = static_cast<double *>(ptr);
  }
  *out = g_scratch[dev];
  return PDA_OK;
}

}  // namespace
}  // namespace pda

using namespace pda;

extern "C" {

int pda_apply_jacobian_noalloc(pda_problem p, const double *d_U, double t, const double *d_B,
                               int64_t ncols, int layout, double *d_AJ, pda_stream stream) {
  if (!p || !d_U || !d_B || !d_AJ || ncols < 0 || (layout != 0 && layout != 1)) {
    set_error("pda_apply_jacobian_noalloc: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  if (p->nnz > kScratchMax) {
    set_error("pda_apply_jacobian_noalloc: the Jacobian has " + std::to_string(p->nnz) +
              " nonzeros, more than the " + std::to_string(kScratchMax) +
              " this entry point holds; use pda_apply_jacobian");
    return PDA_E_INVALID_ARG;
  }
  // the problem's index tables and scalar CSR pattern are uploaded by the
  // allocating entry points on first use; this one only reads them
  if (!p->on_device || !p->d_rowptr.ptr) {
    set_error("pda_apply_jacobian_noalloc: the problem's device tables are not built; call "
              "pda_jacobian_pattern (or pda_apply_jacobian) once before");
    return PDA_E_STATE;
  }
  double *J = nullptr;
  int rc = scratch_on_current_device(&J);
  if (rc) return rc;
  // every value of the pattern is written by the assembly (a fill kernel
  // clears J first where rows accumulate into merged columns), so nothing
  // left in the array by an earlier call is read
  rc = launch_jacobian(*p, d_U, t, J, nullptr, hipStream_t(stream));
  if (rc) return rc;
  return launch_apply(*p, J, d_B, ncols, layout, d_AJ, hipStream_t(stream));
}

}  // extern "C"
