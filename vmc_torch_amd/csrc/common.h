// Shared device helpers and host-side error plumbing for libvmcx.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace vmcx {

// ------------------------------------------------------------ host errors
void set_error(const std::string &msg);

#define VMCX_CHECK_ARG(cond, msg)                                             \
  do {                                                                       \
    if (!(cond)) {                                                           \
      ::vmcx::set_error(std::string(__func__) + ": " + (msg));               \
      return 1;                                                              \
    }                                                                        \
  } while (0)

#define VMCX_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::vmcx::set_error(std::string(__func__) + ": " #call " -> " +          \
                        hipGetErrorString(e_));                              \
      return 2;                                                              \
    }                                                                        \
  } while (0)

#define VMCX_LAUNCHED()                                                       \
  do {                                                                       \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess) {                                                  \
      ::vmcx::set_error(std::string(__func__) + ": launch -> " +             \
                        hipGetErrorString(e_));                              \
      return 2;                                                              \
    }                                                                        \
  } while (0)

inline int nfp_of(int n_f) {
  if (n_f <= 8) return 8;
  if (n_f <= 16) return 16;
  if (n_f <= 32) return 32;
  if (n_f <= 64) return 64;
  return -1;
}

// Largest dynamic LDS a workgroup may request on the current device.
size_t max_lds_per_block();

// Launch-shape table filled by the autotuner (vmcx_set_launch_waves): waves
// per workgroup of a kernel family for a key (its template width and
// precision); `dflt` when nothing was set.
enum { TUNE_SLATER_SWEEP = 0, TUNE_SLATER_ELOC = 1, TUNE_SLATER_REFRESH = 2, TUNE_KINDS = 3 };
int tuned_waves(int kind, int key, int dflt);

// ----------------------------------------------------------- device side
// Philox4x32-10 (Salmon et al., SC'11): 10 rounds, multipliers
// 0xD2511F53 / 0xCD9E8D57, Weyl key increments 0x9E3779B9 / 0xBB67AE85.
struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(uint32_t c0, uint32_t c1,
                                               uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  return {c0, c1, c2, c3};
}

__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
  const uint64_t x = ((uint64_t)hi << 32) | lo;
  return (double)(x >> 11) * 0x1.0p-53;
}

// Bits of sites strictly between a and b (a != b), i.e. (min, max).
__device__ __forceinline__ uint64_t between_mask(int a, int b) {
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  const uint64_t below_hi = (hi >= 64) ? ~0ull : ((1ull << hi) - 1ull);
  const uint64_t upto_lo = (lo >= 63) ? ~0ull : ((1ull << (lo + 1)) - 1ull);
  return below_hi & ~upto_lo;
}

}  // namespace vmcx
