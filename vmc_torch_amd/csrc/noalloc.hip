// Allocation-free variant of vmcx_sr_minres (declared in include/vmcx.h):
// the MINRES workspace lives in a fixed device array of this library instead
// of a caller-supplied buffer, for callers that cannot allocate device memory.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/vmcx.h"
#include "kernels.h"

namespace vmcx {

// The workspace of vmcx_sr_minres_noalloc: a fixed device array, one copy per
// GPU, held until the process ends.  Host code never passes `scratch` itself
// to a kernel (in host code the name is a host-side stand-in); it passes the
// device address that hipGetSymbolAddress returns for the current device.
constexpr int64_t kMinresScratchBytes = (int64_t)20 << 20;  // 20 MiB
static __device__ __align__(256) unsigned char scratch[kMinresScratchBytes];

static void *minres_scratch_ptr() {
  static std::mutex mu;
  static std::vector<void *> per_dev;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if ((int)per_dev.size() <= dev) per_dev.resize(dev + 1, nullptr);
  if (per_dev[dev] == nullptr) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr)
      return nullptr;
    per_dev[dev] = p;
  }
  return per_dev[dev];
}

// sr_minres with its workspace in `scratch`.  Every kernel of sr_minres writes
// the workspace entries it reads earlier in the same call (the scalar block
// is zeroed first, the vectors are copied or zeroed, every partial-sum slot
// is written by its pass), so what a previous call left there is never read.
int sr_minres_noalloc(int64_t Ns, int64_t Np, const void *O, int o_dtype, int64_t ldo,
                      const double *E, double energy_mean, double diag_shift, double rtol,
                      int maxiter, double *dp, double *stats, int32_t *info, hipStream_t st) {
  VMCX_CHECK_ARG(Ns >= 1 && Np >= 1, "Ns and Np must be positive");
  const int64_t need = sr_minres_workspace_bytes(Ns, Np);
  VMCX_CHECK_ARG(need > 0 && need <= kMinresScratchBytes,
                 "Ns / Np too large for the fixed 20 MiB scratch (needs " +
                     std::to_string(need) + " bytes; e.g. Np <= 16384 with Ns <= 65536)");
  void *ws = minres_scratch_ptr();
  VMCX_CHECK_ARG(ws != nullptr, "hipGetSymbolAddress of the MINRES scratch failed");
  return sr_minres(Ns, Np, O, o_dtype, ldo, E, energy_mean, diag_shift, rtol, maxiter, ws,
                   kMinresScratchBytes, dp, stats, info, st);
}

int64_t sr_minres_noalloc_max_bytes() { return kMinresScratchBytes; }

}  // namespace vmcx

// libvmcx status code of an internal return value (0 ok, 1 argument, else HIP)
static inline int noalloc_status(int rc) {
  return rc == 0 ? VMCX_OK : (rc == 1 ? VMCX_EINVAL : VMCX_EHIP);
}

extern "C" {

int64_t vmcx_sr_minres_noalloc_max_bytes(void) { return vmcx::sr_minres_noalloc_max_bytes(); }

int vmcx_sr_minres_noalloc(int64_t Ns, int64_t Np, const void *O, int o_dtype, int64_t ldo,
                           const double *E, double energy_mean, double diag_shift, double rtol,
                           int maxiter, double *dp, double *stats, int32_t *info,
                           void *stream) {
  return noalloc_status(vmcx::sr_minres_noalloc(Ns, Np, O, o_dtype, ldo, E, energy_mean,
                                                diag_shift, rtol, maxiter, dp, stats, info,
                                                (hipStream_t)stream));
}

}
