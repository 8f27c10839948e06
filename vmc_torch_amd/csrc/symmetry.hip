// Point-group images of walker configurations and their fermionic signs
// (the integer part of the symmetry-projected wavefunction,
// GPU/models/symmetry.py:660-770 compute_fermion_signs and the x[:, perms]
// gather of :893-951).
//
// One thread per (configuration b, group element g):
//   images[b][g][s] = x[b][perm[g][s]];
//   sign[b][g] = (-1)^(inversions of q among the occupied modes), modes
//   interleaved m = 2 s + (up ? 1 : 0), q(m) = 2 inv[g][s] + (up ? 1 : 0)
//   with inv[g] the inverse site permutation.  The modes are visited in
//   ascending m; a 128-bit mask holds the q values seen so far, and each
//   new occupied mode adds the number of larger q already seen (popcount).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace vmcx {

__global__ void __launch_bounds__(256)
    sym_images_kernel(int64_t B, int N, int G, const int64_t *__restrict__ cfg,
                      const int32_t *__restrict__ perm, const int32_t *__restrict__ inv,
                      int64_t *images, double *sign) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * G) return;
  const int64_t b = idx / G;
  const int g = (int)(idx - b * G);
  const int64_t *x = cfg + b * N;
  const int32_t *pg = perm + (int64_t)g * N, *ig = inv + (int64_t)g * N;
  int64_t *out = images + idx * N;
  for (int s = 0; s < N; ++s) out[s] = x[pg[s]];
  uint64_t seen_lo = 0, seen_hi = 0;
  int parity = 0;
  for (int s = 0; s < N; ++s) {
    const int v = (int)x[s];
    for (int spin = 0; spin < 2; ++spin) {          // down (bit 0) then up (bit 1)
      if (!((v >> spin) & 1)) continue;
      const int q = 2 * ig[s] + spin;
      // number of seen values greater than q
      const uint64_t above_lo = q >= 63 ? 0ull : (~0ull << (q + 1));
      const uint64_t above_hi = q < 64 ? ~0ull : (q >= 127 ? 0ull : (~0ull << (q - 63)));
      parity += __popcll(seen_lo & above_lo) + __popcll(seen_hi & above_hi);
      if (q < 64)
        seen_lo |= 1ull << q;
      else
        seen_hi |= 1ull << (q - 64);
    }
  }
  sign[idx] = (parity & 1) ? -1.0 : 1.0;
}

int sym_images(int64_t B, int N, int G, const int64_t *cfg, const int32_t *perm,
               const int32_t *inv, int64_t *images, double *sign, hipStream_t st) {
  VMCX_CHECK_ARG(B >= 0 && N >= 1 && N <= 64 && G >= 1 && G <= 64, "bad sizes");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(cfg && perm && inv && images && sign, "null pointer");
  const int64_t n = B * G;
  hipLaunchKernelGGL(sym_images_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, B,
                     N, G, cfg, perm, inv, images, sign);
  VMCX_LAUNCHED();
  return 0;
}

}  // namespace vmcx
