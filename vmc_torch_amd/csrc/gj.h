// Register-resident Gauss-Jordan inversion shared by the determinant kernels.
#pragma once

#include "common.h"

namespace vmcx {

__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct GJResult {
  double logdet;
  int parity;
  bool singular;
  int my_step;
};

// In-place Gauss-Jordan with partial pivoting on B = A^T held row-per-lane
// within a group of NFP lanes (lane k = row k of B = column k of A, register
// j = row j of A).  Pivot rows are lanes; the pivot column is always register
// 0 because the register file is rotated by one after each step (NFP
// rotations restore the order).  On exit lane q_p holds column p of A^-1
// with register c = row q_c (q stored in qbuf), so the caller scatters
// register c to ainv[p][qbuf[c]].  det(A) = (-1)^parity * exp(logdet).
// prow / qbuf: NFP doubles / NFP bytes of LDS private to the group.
template <int NFP>
__device__ __forceinline__ GJResult gj_invert(double (&a)[NFP], int col, int gbase,
                                             double *prow, uint8_t *qbuf) {
  double logdet = 0.0;
  int parity = 0;
  uint64_t used = 0;
  int my_step = -1;
  bool singular = false;
#pragma unroll 1
  for (int p = 0; p < NFP; ++p) {
    // NaN / Inf entries rank as 0 so that the order stays total and every
    // lane is chosen exactly once over the NFP steps
    const double mag = fabs(a[0]);
    double cand = ((used >> col) & 1ull) ? -1.0 : ((mag <= 1e300) ? mag : 0.0);
    int cidx = col;
#pragma unroll
    for (int o = NFP / 2; o > 0; o >>= 1) {
      const double oc = __shfl_xor(cand, o, 64);
      const int oi = __shfl_xor(cidx, o, 64);
      if (oc > cand || (oc == cand && oi < cidx)) {
        cand = oc;
        cidx = oi;
      }
    }
    const int q = cidx;
    if (!(cand > 0.0)) singular = true;
    const double piv = __shfl(a[0], gbase + q, 64);
    const double f = a[0];
    if (col == q) {
      const double inv = (piv != 0.0) ? 1.0 / piv : 0.0;
#pragma unroll
      for (int j = 1; j < NFP; ++j) a[j] *= inv;
      a[0] = inv;
#pragma unroll
      for (int j = 0; j < NFP; ++j) prow[j] = a[j];
      qbuf[p] = (uint8_t)q;
      my_step = p;
    }
    wave_lds_fence();
    if (col != q) {
      a[0] = 0.0;
#pragma unroll
      for (int j = 0; j < NFP; ++j) a[j] = fma(-f, prow[j], a[j]);
    }
    wave_lds_fence();
    logdet += (piv != 0.0) ? log(fabs(piv)) : 0.0;
    parity += (piv < 0.0) ? 1 : 0;
    parity += __popcll(used & ~((2ull << q) - 1ull));
    used |= 1ull << q;
    const double a0 = a[0];
#pragma unroll
    for (int j = 0; j < NFP - 1; ++j) a[j] = a[j + 1];
    a[NFP - 1] = a0;
  }
  wave_lds_fence();
  return {logdet, parity, singular, my_step};
}

}  // namespace vmcx
