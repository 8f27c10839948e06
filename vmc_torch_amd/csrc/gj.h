// Register-resident Gauss-Jordan inversion shared by the determinant kernels.
#pragma once

#include "common.h"

namespace vmcx {

__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// wave-wide helpers for one 64 x 64 matrix per wave: a double read from
// lane l into scalar registers, and the maximum of a 32-bit key over the
// wave by a DPP max-scan (row_shr 1, 2, 4, 8, row_bcast 15, 31: the maximum
// ends in lane 63)
__device__ __forceinline__ double lu_readlane(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ uint32_t lu_wave_max(uint32_t k) {
  k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x111, 0xf, 0xf, false));
  k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x112, 0xf, 0xf, false));
  k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x114, 0xf, 0xf, false));
  k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x118, 0xf, 0xf, false));
  k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x142, 0xa, 0xf, false));
  k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x143, 0xc, 0xf, false));
  return (uint32_t)__builtin_amdgcn_readlane((int)k, 63);
}

// two 32-lane groups per wave: the maximum of each half (row_shr 1, 2, 4,
// 8, row_bcast 15 into rows 1 and 3 leaves it in lanes 31 and 63), and
// the pivot value of each half read from its group
__device__ __forceinline__ void lu_half_pivot(uint32_t k, double v, int lane, int &q,
                                              double &piv) {
  k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x111, 0xf, 0xf, false));
  k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x112, 0xf, 0xf, false));
  k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x114, 0xf, 0xf, false));
  k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x118, 0xf, 0xf, false));
  k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x142, 0xa, 0xf, false));
  const int q0 = 63 - (int)((uint32_t)__builtin_amdgcn_readlane((int)k, 31) & 63u);
  const int q1 = 63 - (int)((uint32_t)__builtin_amdgcn_readlane((int)k, 63) & 63u);
  const double p0 = lu_readlane(v, q0), p1 = lu_readlane(v, 32 + q1);
  q = lane < 32 ? q0 : q1;
  piv = lane < 32 ? p0 : p1;
}

// True when an entry of the group's matrix (lanes gbase .. gbase + NFP - 1)
// is NaN, infinite or beyond 1e300: the determinant is then reported as NaN
// (a non-finite parameter must not pass for a zero amplitude).
// largest magnitude an LU entry may have (larger ones rank as non-finite)
template <typename T>
__device__ __forceinline__ T lu_big();
template <>
__device__ __forceinline__ double lu_big<double>() { return 1e300; }
template <>
__device__ __forceinline__ float lu_big<float>() { return 3.0e38f; }

// lane l's value on every lane (readlane into scalar registers)
__device__ __forceinline__ double lu_bcast(double v, int l) { return lu_readlane(v, l); }
__device__ __forceinline__ float lu_bcast(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

template <int NFP, typename T>
__device__ __forceinline__ bool group_nonfinite(const T (&a)[NFP], int gbase) {
  bool bad = false;
#pragma unroll
  for (int j = 0; j < NFP; ++j) bad |= !(fabs(a[j]) <= lu_big<T>());
  const uint64_t b = __ballot(bad);
  const uint64_t m = (NFP == 64) ? ~0ull : (((1ull << NFP) - 1ull) << gbase);
  return (b & m) != 0ull;
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
    int q;
    double piv;
    if constexpr (NFP == 64 || NFP == 32) {
      // one or two matrices per wave: DPP arg-max on a 32-bit key (|a| as
      // float, the low 6 bits 63 - lane), pivot value by readlane
      const float fm = (mag <= 1e300) ? (float)mag : 0.0f;
      const uint32_t key = ((used >> col) & 1ull)
                               ? 0u
                               : ((__float_as_uint(fm) & ~63u) | (uint32_t)(63 - col));
      if constexpr (NFP == 64) {
        q = 63 - (int)(lu_wave_max(key) & 63u);
        piv = lu_readlane(a[0], q);
      } else {
        lu_half_pivot(key, a[0], gbase + col, q, piv);
      }
      if (!(piv != 0.0) || !(fabs(piv) <= 1e300)) singular = true;
    } else {
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
      q = cidx;
      if (!(cand > 0.0)) singular = true;
      piv = __shfl(a[0], gbase + q, 64);
    }
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

// Determinant only: LU with the same pivot rule on the same layout (lane k =
// row k of B = A^T, register j = column j of B), eliminating below the pivot
// only.  Each step is instantiated with its index as a template argument, so
// step p touches registers p+1.. only (half the FMAs of Gauss-Jordan) and
// the register array is never indexed at run time; |det| is accumulated as
// a mantissa product and an exponent sum (frexp), one log per determinant.
struct LuState {
  double mprod;
  int esum, parity;
  uint64_t used;
  bool singular;
};

template <int NFP, int P, typename T>
struct LuStep {
  static __device__ __forceinline__ void run(T (&a)[NFP], int col, int gbase, T *prow,
                                             LuState &S) {
    // pivot key: |a| rounded to float with its 6 low mantissa bits replaced
    // by 63 - lane (largest magnitude to 2^-17, then the lowest lane); used
    // lanes key 0.  NaN / Inf rank as 0 as in gj_invert.
    const bool done = (S.used >> col) & 1ull;
    const T mag = fabs(a[P]);
    const float fm = (mag <= lu_big<T>()) ? (float)mag : 0.0f;
    uint32_t key = done ? 0u : ((__float_as_uint(fm) & ~63u) | (uint32_t)(63 - col));
    int q;
    T piv;
    if constexpr (NFP == 32) {
      if constexpr (sizeof(T) == 8) {
        lu_half_pivot(key, a[P], gbase + col, q, piv);
      } else {
        double pd;
        lu_half_pivot(key, (double)a[P], gbase + col, q, pd);
        piv = (T)pd;
      }
    } else {
#pragma unroll
      for (int o = NFP / 2; o > 0; o >>= 1)
        key = max(key, (uint32_t)__shfl_xor((int)key, o, 64));
      q = 63 - (int)(key & 63u);
      piv = __shfl(a[P], gbase + q, 64);
    }
    if (!(piv != T(0)) || !(fabs(piv) <= lu_big<T>())) S.singular = true;
    if (P + 1 < NFP) {
      if (col == q) {
#pragma unroll
        for (int j = P + 1; j < NFP; ++j) prow[j] = a[j];
      }
      wave_lds_fence();
      if (!done && col != q) {
        const T f = (piv != T(0)) ? a[P] / piv : T(0);
        // blocks of 16 pivot-row entries, kept apart by scheduling barriers
        // so that at most 16 of them are live (no spills at NFP = 64)
#pragma unroll
        for (int j0 = P + 1; j0 < NFP; j0 += 16) {
#pragma unroll
          for (int j = j0; j < j0 + 16 && j < NFP; ++j) a[j] = fma(-f, prow[j], a[j]);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      wave_lds_fence();
    }
    if (piv != T(0)) {
      int e;
      S.mprod *= frexp(fabs((double)piv), &e);
      S.esum += e;
      S.mprod = frexp(S.mprod, &e);
      S.esum += e;
    }
    S.parity += (piv < T(0)) ? 1 : 0;
    S.parity += __popcll(S.used & ~((2ull << q) - 1ull));
    S.used |= 1ull << q;
    LuStep<NFP, P + 1, T>::run(a, col, gbase, prow, S);
  }
};

template <int NFP, typename T>
struct LuStep<NFP, NFP, T> {
  static __device__ __forceinline__ void run(T (&)[NFP], int, int, T *, LuState &) {}
};

// One 64 x 64 matrix per wave: the pivot is the same for every lane, so the
// arg-max is a DPP max-scan (row_shr 1, 2, 4, 8, row_bcast 15, 31; the wave
// maximum ends in lane 63).  The pivot row reaches the other lanes two ways
// at once: its head by readlane into scalar registers (two VALU
// instructions per entry with the update), its tail through the wave's
// 64-entry LDS buffer (written by the pivot lane as 16-byte vectors, read
// back by every lane as broadcasts: one VALU instruction per entry); the
// tail takes 4/6 of the entries in FP32, 3/6 in FP64, the split that keeps
// both the VALU and the LDS pipes busy.  Every lane applies the update, the
// pivot lane with multiplier exactly 1: its registers past P become exact
// zeros, so a lane once chosen keys at most 63 (the lane bits alone) at
// every later step and loses to any unchosen lane with a nonzero entry; no
// per-lane "done" test is needed.  (If every unchosen entry is zero the
// matrix is singular and whichever lane wins carries a zero pivot.)  The
// pivot of each step is kept in the chosen lane of `pivs`; sign and
// log|det| are formed from them once, after the last step (lu64_finish).
// multiplier a / piv: FP32 by the hardware reciprocal (1 ulp), FP64 by the
// correctly rounded division
__device__ __forceinline__ float lu_mult(float a, float piv) {
  return a * __builtin_amdgcn_rcpf(piv);
}
__device__ __forceinline__ double lu_mult(double a, double piv) { return a / piv; }

// 16-byte vectors of the LDS part
template <typename T>
struct LuVec;
template <>
struct LuVec<float> {
  typedef float4 t;
  static constexpr int n = 4, tail6 = 4;
  static __device__ __forceinline__ t make(const float *x) {
    return make_float4(x[0], x[1], x[2], x[3]);
  }
};
template <>
struct LuVec<double> {
  typedef double2 t;
  static constexpr int n = 2, tail6 = 3;
  static __device__ __forceinline__ t make(const double *x) { return make_double2(x[0], x[1]); }
};

template <int P, typename T>
struct LuStep64 {
  static __device__ __forceinline__ void run(T (&a)[64], int col, T &pivs, uint64_t &used,
                                             int &parity, T *buf) {
    // pivot key: |a| rounded to float with its 6 low mantissa bits replaced
    // by 63 - lane (largest magnitude to 2^-17, then the lowest lane); NaN /
    // Inf rank as 0
    const T mag = fabs(a[P]);
    const float fm = (mag <= lu_big<T>()) ? (float)mag : 0.0f;
    const uint32_t key = (__float_as_uint(fm) & ~63u) | (uint32_t)(63 - col);
    const int q = 63 - (int)(lu_wave_max(key) & 63u);
    const T piv = lu_bcast(a[P], q);
    const bool mine = col == q;
    pivs = mine ? piv : pivs;
    if (P + 1 < 64) {
      const T f = mine ? T(1) : lu_mult(a[P], piv);
      typedef LuVec<T> V;
      // entries [P + 1, js) by readlane, [js, 64) through LDS (js aligned)
      constexpr int n_tail = ((63 - P) * V::tail6) / 6;
      constexpr int js0 = (64 - n_tail + V::n - 1) & ~(V::n - 1);
      constexpr int js = js0 < P + 1 ? ((P + 1 + V::n - 1) & ~(V::n - 1)) : js0;
      if (js < 64 && mine) {
#pragma unroll
        for (int j = js; j < 64; j += V::n)
          *reinterpret_cast<typename V::t *>(buf + j) = V::make(&a[j]);
      }
      // blocks of 8 readlane entries kept apart by scheduling barriers, so
      // that the readlanes of one block are consumed before the next block's
      // are issued (hoisting all of them overflows the scalar registers)
#pragma unroll
      for (int j0 = P + 1; j0 < js && j0 < 64; j0 += 8) {
        T r[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (j0 + k < js && j0 + k < 64) r[k] = lu_bcast(a[j0 + k], q);
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (j0 + k < js && j0 + k < 64) a[j0 + k] = fma(-f, r[k], a[j0 + k]);
        __builtin_amdgcn_sched_barrier(0);
      }
      if (js < 64) {
        wave_lds_fence();
#pragma unroll
        for (int j = js; j < 64; j += V::n) {
          const typename V::t v = *reinterpret_cast<const typename V::t *>(buf + j);
          a[j] = fma(-f, v.x, a[j]);
          a[j + 1] = fma(-f, v.y, a[j + 1]);
          if constexpr (V::n == 4) {
            a[j + 2] = fma(-f, v.z, a[j + 2]);
            a[j + 3] = fma(-f, v.w, a[j + 3]);
          }
        }
        // the buffer is rewritten by the next step's pivot lane
        wave_lds_fence();
      }
    }
    // row exchanges: lanes chosen earlier with a higher index than q
    parity += __popcll(used & ~((2ull << q) - 1ull));
    used |= 1ull << q;
    LuStep64<P + 1, T>::run(a, col, pivs, used, parity, buf);
  }
};

template <typename T>
struct LuStep64<64, T> {
  static __device__ __forceinline__ void run(T (&)[64], int, T &, uint64_t &, int &, T *) {}
};

// sign and log|det| from the 64 pivots (lane l: the pivot taken from lane
// l), summed in FP64 over the wave; a zero, non-finite or missing pivot
// (a lane never chosen keeps 0) makes the matrix singular
template <typename T>
__device__ __forceinline__ void lu64_finish(T pivs, uint64_t used, int parity, double &logdet,
                                            int &par, bool &singular) {
  const bool bad = !(pivs != T(0)) || !(fabs(pivs) <= lu_big<T>());
  singular = (__ballot(bad) != 0ull) || (used != ~0ull);
  par = parity + __popcll(__ballot(pivs < T(0)));
  double l = bad ? 0.0 : log(fabs((double)pivs));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  logdet = l;
}

// T = double, or float (the mixed-precision amplitude: FP32 elimination,
// log|det| still accumulated in FP64: as a sum of the pivots' logarithms at
// NFP = 64, as a mantissa product and an exponent sum otherwise)
template <int NFP, typename T>
__device__ __forceinline__ GJResult lu_logdet(T (&a)[NFP], int col, int gbase, T *prow) {
  if constexpr (NFP == 64) {
    T pivs = T(0);
    uint64_t used = 0ull;
    int parity = 0;
    LuStep64<0, T>::run(a, col, pivs, used, parity, prow);
    GJResult r{0.0, 0, false, -1};
    lu64_finish(pivs, used, parity, r.logdet, r.parity, r.singular);
    return r;
  } else {
    LuState S{1.0, 0, 0, 0ull, false};
    LuStep<NFP, 0, T>::run(a, col, gbase, prow, S);
    return {log(S.mprod) + (double)S.esum * 0.69314718055994530942, S.parity, S.singular, -1};
  }
}

}  // namespace vmcx
