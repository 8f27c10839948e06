// Blocked FP64 Cholesky factorisation and triangular solves for the SR
// system (column-major n x n, lower triangle of A as input).
//
// Factorisation, right-looking with block NB = 128:
//   for each diagonal block k:
//     chol_diag_kernel  : L_kk = chol(A_kk) in LDS (one workgroup), and
//                         Dinv_k = L_kk^-1 (kept for the solves; L_kk itself
//                         is never written back: every later use of the
//                         diagonal block goes through Dinv_k)
//     GEMM with Dinv_k  : W_k = A_k,rest L_kk^-T        (panel below, into W)
//     GEMM / SYRK       : A_rest <- A_rest - W W^T      (trailing update)
// The factor is Dinv (diagonal blocks, inverted) plus W (off-diagonal
// blocks); A is left as scratch.  The diagonal block stays in LDS for all
// NB steps of its factorisation and inversion (one launch per block).
//
// Solves use the inverted diagonal blocks so that each block step is one
// launch: forward  y_k = Dinv_k (b_k) then b_rest -= L_rest,k y_k;
//         backward x_k = Dinv_k^T (y_k) then y_head -= L_k,head^T x_k.
#include <rocblas/rocblas.h>

#include <mutex>
#include <map>
#include <unordered_map>
#include <vector>

#include "kernels.h"

namespace vmcx {

constexpr int CHOL_NB = 128;

// A: column-major, element (i, j) at A[j*lda + i].
// Dinv_k (column-major NB x NB): Dinv[m*NB + c] = (L_kk^-1)[c][m].
// DinvT_k (column-major):       DinvT[c*NB + m] = (L_kk^-1)[c][m].
//
// The 128 x 128 diagonal block is factorised in LDS in 16-column steps t:
//   leaf     : one wave factorises and inverts the 16 x 16 diagonal tile
//              as 2 x 2 blocks of 8 x 8 (each 8 x 8 Cholesky + inverse in
//              one lane's registers, the coupling blocks lane-parallel);
//              L_tt stays in the lower triangle of the tile, the strictly
//              lower part of X_tt = L_tt^-1 goes transposed into its
//              (unused) upper triangle, and the reciprocal diagonal
//              1 / L_ii (= X_ii) into rdiag
//   panel    : L_I,t = A_I,t X_tt^T for every tile I below (matrix cores)
//   trailing : C_IJ -= L_I,t L_J,t^T for t < J <= I (matrix cores); the wave
//              that updates tile (t+1, t+1) factorises it at once, while the
//              other waves finish the rest of the trailing update
// The inverse X = L^-1 of the whole block is then formed in place, tile
// column jt from the right:  X_T,jt = -(sum_{jt < K <= T} X_TK L_K,jt) X_jt,jt.
//
// v_mfma_f64_16x16x4f64 operand layout (lane l of the wave, k-step of 4):
//   A (16 x 4): lane l holds A[l % 16][l / 16]
//   B (4 x 16): lane l holds B[l / 16][l % 16]
//   D (16 x 16, four FP64 registers r): lane l, r holds D[4 r + l / 16][l % 16]
constexpr int LEAF = 16;
constexpr int DIAG_THREADS = 512;
typedef double dbl4 __attribute__((ext_vector_type(4)));

// 1/sqrt(d) for d > 0: hardware estimate and two Newton steps
__device__ __forceinline__ double rsqrt_f64(double d) {
  double y = __builtin_amdgcn_rsq(d);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double h = d * y;
    const double e = fma(-h, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 8x8 Cholesky and inverse of the lower triangle stored packed in a[]
// (a[i(i+1)/2 + j], j <= i), in one lane's registers: a becomes L, rs[i] =
// 1 / L_ii, and X = L^-1 is formed one column at a time (eight live values)
// and stored transposed through T: T[o + c][o + i] = X[i][c] for i > c.
// Returns the first non-positive pivot (1-based) or 0.
__device__ __forceinline__ int chol8_inv8(double *a, double *rs, double (*T)[LEAF + 1], int o) {
  int badj = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int jj = j * (j + 1) / 2 + j;
    const double d = a[jj];
    const bool ok = d > 0.0;
    if (!ok && badj == 0) badj = j + 1;
    const double r = ok ? rsqrt_f64(d) : 1.0;
    rs[j] = r;
    a[jj] = ok ? d * r : 1.0;
#pragma unroll
    for (int i = j + 1; i < 8; ++i) a[i * (i + 1) / 2 + j] *= r;
#pragma unroll
    for (int i = j + 1; i < 8; ++i)
#pragma unroll
      for (int l = j + 1; l <= i; ++l)
        a[i * (i + 1) / 2 + l] = fma(-a[i * (i + 1) / 2 + j], a[l * (l + 1) / 2 + j],
                                     a[i * (i + 1) / 2 + l]);
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    double x[8];
    x[c] = rs[c];
#pragma unroll
    for (int i = c + 1; i < 8; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int q = c; q < i; ++q) acc = fma(a[i * (i + 1) / 2 + q], x[q], acc);
      x[i] = -rs[i] * acc;
      T[o + c][o + i] = x[i];
    }
  }
  return badj;
}

// The 128 x 128 block lives in LDS as its 36 lower 16 x 16 tiles (rows padded
// to 17): 78 KB, so the kernel fits on a CU beside one workgroup of the
// trailing-update SYRK (74 KB) instead of waiting for a whole CU.  Tile
// (I, J), I >= J, is Lt[I (I + 1) / 2 + J]; a diagonal tile's upper triangle
// holds the strictly lower part of its leaf inverse, transposed.
constexpr int LTP = LEAF + 1;
typedef double LTile[LEAF][LTP];
constexpr int NLT = (CHOL_NB / LEAF) * (CHOL_NB / LEAF + 1) / 2;
__device__ __forceinline__ double &lt_at(LTile *Lt, int i, int j) {
  const int I = i >> 4, J = j >> 4;
  return Lt[I * (I + 1) / 2 + J][i & (LEAF - 1)][j & (LEAF - 1)];
}
__device__ __forceinline__ const double &lt_at(const LTile *Lt, int i, int j) {
  const int I = i >> 4, J = j >> 4;
  return Lt[I * (I + 1) / 2 + J][i & (LEAF - 1)][j & (LEAF - 1)];
}

// Leaf of step kb (wave 0): factor and invert the 16 x 16 tile at (kb, kb)
// as 2 x 2 blocks of 8 x 8:  L11 = chol(A11), X11 = L11^-1 (one lane);
// L21 = A21 X11^T (8 lanes);  A22 -= L21 L21^T (36 lanes);  L22, X22 (one
// lane);  X21 = -X22 L21 X11 (64 lanes, via Y = L21 X11 in `scr`).  L stays
// in the lower triangle of the tile; the strictly lower part of
// X = L^-1 goes transposed into its upper triangle; rdiag_all = 1 / L_ii.
__device__ __forceinline__ void chol_leaf(LTile *L, double *rdiag_all, int kb,
                                          int lane, int *bad, double *scr) {
  double (*T)[LTP] = L[(kb >> 4) * ((kb >> 4) + 1) / 2 + (kb >> 4)];
  // ---- 1: L11, X11
  if (lane == 0) {
    double a[36], rs[8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) a[i * (i + 1) / 2 + j] = T[i][j];
    const int bj = chol8_inv8(a, rs, T, 0);
    if (bj && *bad == 0) *bad = kb + bj;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      rdiag_all[kb + i] = rs[i];
#pragma unroll
      for (int j = 0; j <= i; ++j) T[i][j] = a[i * (i + 1) / 2 + j];
    }
  }
  wave_lds_sync();
  // ---- 2: L21[r][c] = sum_{q <= c} A21[r][q] X11[c][q]  (lane r - 8)
  if (lane < 8) {
    const int r = 8 + lane;
    double arow[8], out[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) arow[q] = T[r][q];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      double acc = arow[c] * rdiag_all[kb + c];
#pragma unroll
      for (int q = 0; q < c; ++q) acc = fma(arow[q], T[q][c], acc);   // X11[c][q] at T[q][c]
      out[c] = acc;
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) T[r][c] = out[c];
  }
  wave_lds_sync();
  // ---- 3: A22[i][l] -= L21[i] . L21[l]  (i >= l, packed over 36 lanes)
  if (lane < 36) {
    int i = 0;
    while ((i + 1) * (i + 2) / 2 <= lane) ++i;
    const int l = lane - i * (i + 1) / 2;
    double acc = T[8 + i][8 + l];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc = fma(-T[8 + i][c], T[8 + l][c], acc);
    T[8 + i][8 + l] = acc;
  }
  wave_lds_sync();
  // ---- 4: L22, X22
  if (lane == 0) {
    double a[36], rs[8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) a[i * (i + 1) / 2 + j] = T[8 + i][8 + j];
    const int bj = chol8_inv8(a, rs, T, 8);
    if (bj && *bad == 0) *bad = kb + 8 + bj;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      rdiag_all[kb + 8 + i] = rs[i];
#pragma unroll
      for (int j = 0; j <= i; ++j) T[8 + i][8 + j] = a[i * (i + 1) / 2 + j];
    }
  }
  wave_lds_sync();
  // ---- 5: Y = L21 X11 (Y[i][c] = sum_{q >= c} L21[i][q] X11[q][c]), then
  //         X21 = -X22 Y (X21[i][c] = -sum_{q <= i} X22[i][q] Y[q][c])
  {
    const int i = lane >> 3, c = lane & 7;
    double acc = T[8 + i][c] * rdiag_all[kb + c];
#pragma unroll
    for (int q = c + 1; q < 8; ++q) acc = fma(T[8 + i][q], T[c][q], acc);  // X11[q][c] at T[c][q]
    scr[lane] = acc;
  }
  wave_lds_sync();
  {
    const int i = lane >> 3, c = lane & 7;
    double acc = rdiag_all[kb + 8 + i] * scr[i * 8 + c];
    for (int q = 0; q < i; ++q) acc = fma(T[8 + q][8 + i], scr[q * 8 + c], acc);  // X22[i][q]
    T[c][8 + i] = -acc;   // X[8+i][c], stored transposed
  }
}

// X_tt (lower, the leaf inverse of tile t at kb) as a matrix-core B operand:
// element [k][j] of X_tt^T (transpose = 0) or of X_tt (transpose = 1)
__device__ __forceinline__ double leaf_inv(const LTile *L,
                                          const double *rdiag_all, int kb, int row, int col) {
  // X_tt[row][col]
  if (row == col) return rdiag_all[kb + row];
  return (row > col) ? lt_at(L, kb + col, kb + row) : 0.0;
}

__global__ void __launch_bounds__(DIAG_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4)))
    chol_diag_kernel(double *A, int64_t lda, int64_t k0, int nbk, int32_t *info,
                     double *Dinv, int32_t *busy) {
  // the factorisation's critical path: its waves go first when the CU is
  // shared with trailing-update workgroups, and `busy` names its CU to the
  // trailing SYRKs, whose waves on this CU sleep until it is cleared
  __builtin_amdgcn_s_setprio(3);
  if (busy && threadIdx.x == 0)
    __hip_atomic_store(busy, cu_code(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __shared__ LTile L[NLT];                    // lower tiles of the block
  __shared__ double W[64];                    // leaf scratch
  __shared__ double rdiag_all[CHOL_NB];       // 1 / lt_at(L, i, i)
  __shared__ int bad;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  constexpr int NWAVE = DIAG_THREADS / 64;
  constexpr int NT = CHOL_NB / LEAF;
  double *Akk = A + k0 * lda + k0;
  // all loads issued before the first LDS store (16 per thread)
  {
    double v[CHOL_NB * CHOL_NB / DIAG_THREADS];
#pragma unroll
    for (int q = 0; q < CHOL_NB * CHOL_NB / DIAG_THREADS; ++q) {
      const int idx = tid + q * DIAG_THREADS;
      const int j = idx / CHOL_NB, i = idx - j * CHOL_NB;
      double x = 0.0;
      if (i < nbk && j < nbk) x = (i >= j) ? Akk[(int64_t)j * lda + i] : 0.0;
      else if (i == j) x = 1.0;  // identity padding keeps the leaves regular
      v[q] = x;
    }
#pragma unroll
    for (int q = 0; q < CHOL_NB * CHOL_NB / DIAG_THREADS; ++q) {
      const int idx = tid + q * DIAG_THREADS;
      const int j = idx / CHOL_NB, i = idx - j * CHOL_NB;
      if ((i >> 4) >= (j >> 4)) lt_at(L, i, j) = v[q];  // lower tiles only
    }
  }
  if (tid == 0) bad = 0;
  __syncthreads();
  if (wave == 0) chol_leaf(L, rdiag_all, 0, lane, &bad, W);
  __syncthreads();

  for (int t = 0; t < NT; ++t) {
    const int kb = t * LEAF;
    const int nt = NT - 1 - t;  // tiles below the leaf
    if (nt == 0) break;
    // ---- panel: L_I,t = A_I,t X_tt^T, one wave per tile
    if (wave < nt) {
      const int r0 = kb + LEAF * (wave + 1);
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s4 = 0; s4 < LEAF; s4 += 4) {
        const int kk = s4 + lk;
        const double a = lt_at(L, r0 + li, kb + kk);
        const double b = leaf_inv(L, rdiag_all, kb, li, kk);  // X_tt^T[kk][li]
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      }
      wave_lds_sync();
#pragma unroll
      for (int r = 0; r < 4; ++r) lt_at(L, r0 + 4 * r + lk, kb + li) = acc[r];
    }
    __syncthreads();
    // ---- trailing update C_IJ -= P_I P_J^T (1 <= J <= I <= nt, relative
    //      to the leaf); wave 0 takes tile (1, 1) and then factorises it
    {
      const int ntiles = nt * (nt + 1) / 2;
      // tile 0 = (1, 1) to wave 0 alone, the others to waves 1..15
      const int stride = (wave == 0) ? ntiles : NWAVE - 1;
      for (int q = wave; q < ntiles; q += stride) {
        int I = 1, rem = q;
        while (rem >= I) { rem -= I; ++I; }
        const int J = rem + 1;
        const int r0 = kb + LEAF * I, c0 = kb + LEAF * J;
        dbl4 acc;
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = lt_at(L, r0 + 4 * r + lk, c0 + li);
#pragma unroll
        for (int s4 = 0; s4 < LEAF; s4 += 4) {
          const double a = -lt_at(L, r0 + li, kb + s4 + lk);
          const double b = lt_at(L, c0 + li, kb + s4 + lk);
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) lt_at(L, r0 + 4 * r + lk, c0 + li) = acc[r];
      }
      if (wave == 0) {
        wave_lds_sync();
        chol_leaf(L, rdiag_all, kb + LEAF, lane, &bad, W);
      }
    }
    __syncthreads();
  }
  if (bad) {
    if (tid == 0 && info && info[0] == 0 && bad <= nbk) info[0] = (int32_t)(k0 + bad);
    if (busy && tid == 0) __hip_atomic_store(busy, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  // (L_kk itself is not stored: the solves and the panel use its inverse)
  // ---- X = L^-1 in place by recursive doubling (the diagonal tiles already
  //      hold their inverses): for block sizes h = 16, 32, 64 and each pair
  //      (A, B) of consecutive h-blocks,  X_BA = -X_BB (L_BA X_AA).  A level
  //      forms all tiles of P = L_BA X_AA in registers, stores them over L_BA
  //      after a barrier, then does the same for X_BA = -X_BB P.  Levels
  //      h = 16, 32 have 4 and 8 tiles (one per wave); at h = 64 (16 tiles)
  //      a wave takes P(i, j) and P(i, 3 - j), then X(i, j) and X(3 - i, j):
  //      P(i, j) sums 4 - j tile products and X(i, j) sums i + 1, so every
  //      wave does five.
#pragma unroll
  for (int lvl = 0; lvl < 3; ++lvl) {
    const int ht = 1 << lvl;  // tiles per half block
#pragma unroll
    for (int phase = 0; phase < 2; ++phase) {
      int ti[2], tj[2], tA0[2];
      bool on[2];
      if (lvl < 2) {
        const int pair = wave / (ht * ht), r = wave % (ht * ht);
        on[0] = wave < 4 * ht;
        on[1] = false;
        ti[0] = ti[1] = r / ht;
        tj[0] = tj[1] = r % ht;
        tA0[0] = tA0[1] = 2 * ht * pair;
      } else {
        on[0] = on[1] = true;
        tA0[0] = tA0[1] = 0;
        if (phase == 0) {
          ti[0] = ti[1] = wave / 2;
          tj[0] = wave % 2;
          tj[1] = 3 - wave % 2;
        } else {
          tj[0] = tj[1] = wave / 2;
          ti[0] = wave % 2;
          ti[1] = 3 - wave % 2;
        }
      }
      dbl4 acc[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        acc[q] = dbl4{0.0, 0.0, 0.0, 0.0};
        if (!on[q]) continue;
        const int i = ti[q], j = tj[q], A0 = tA0[q], B0 = A0 + ht;
        if (phase == 0) {  // P(i, j) = sum_{k >= j} L_BA(i, k) X_AA(k, j)
          for (int k = j; k < ht; ++k) {
#pragma unroll
            for (int s4 = 0; s4 < LEAF; s4 += 4) {
              const int kk = s4 + lk;
              const double a = lt_at(L, LEAF * (B0 + i) + li, LEAF * (A0 + k) + kk);
              const double b = (k == j) ? leaf_inv(L, rdiag_all, LEAF * (A0 + j), kk, li)
                                        : lt_at(L, LEAF * (A0 + k) + kk, LEAF * (A0 + j) + li);
              acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
            }
          }
        } else {  // X_BA(i, j) = -sum_{k <= i} X_BB(i, k) P(k, j)
          for (int k = 0; k <= i; ++k) {
#pragma unroll
            for (int s4 = 0; s4 < LEAF; s4 += 4) {
              const int kk = s4 + lk;
              const double a = -((k == i) ? leaf_inv(L, rdiag_all, LEAF * (B0 + i), li, kk)
                                          : lt_at(L, LEAF * (B0 + i) + li, LEAF * (B0 + k) + kk));
              const double b = lt_at(L, LEAF * (B0 + k) + kk, LEAF * (A0 + j) + li);
              acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
            }
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (!on[q]) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          lt_at(L, LEAF * (tA0[q] + ht + ti[q]) + 4 * r + lk, LEAF * (tA0[q] + tj[q]) + li) =
              acc[q][r];
      }
      __syncthreads();
    }
  }
  // Dinv[a*NB + b] = X[b][a]  (DinvT, its transpose, is written after the
  // factorisation by chol_dinv_transpose_kernel, off the critical path);
  // unrolled so that the LDS reads of eight stores are in flight together
#pragma unroll 8
  for (int idx = tid; idx < CHOL_NB * CHOL_NB; idx += DIAG_THREADS) {
    const int a = idx / CHOL_NB, b = idx - a * CHOL_NB;
    double xba = 0.0;
    if (a < nbk && b < nbk) {
      if ((a >> 4) == (b >> 4)) {
        const int tb = a & ~15;
        xba = leaf_inv(L, rdiag_all, tb, b - tb, a - tb);
      } else if (b > a) {
        xba = lt_at(L, b, a);
      }
    }
    Dinv[idx] = xba;
  }
  if (busy && tid == 0) __hip_atomic_store(busy, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// DinvT_k = Dinv_k^T for every block (grid: blocks x 64)
__global__ void __launch_bounds__(256)
    chol_dinv_transpose_kernel(int64_t nblk, const double *Dinv, double *DinvT) {
  const int64_t kb = blockIdx.x / 64;
  const int part = blockIdx.x % 64;
  if (kb >= nblk) return;
  const double *S = Dinv + kb * CHOL_NB * CHOL_NB;
  double *D = DinvT + kb * CHOL_NB * CHOL_NB;
  const int idx = part * 256 + threadIdx.x;          // 64 x 256 = NB^2
  const int a = idx / CHOL_NB, b = idx - a * CHOL_NB;
  D[idx] = S[b * CHOL_NB + a];
}

// ---------------------------------------------------------------- solves
// Each solve is a chain of one launch per diagonal block, so the launches are
// built for latency.  Workgroups of 1024 threads hold a 128 x 128 block as
// 128 rows x 8 groups of 16, so every thread issues its 16 (+16) loads at once
// (one memory round trip), and workgroup 0 of each launch also forms the
// next block's solution, so no launch repeats a diagonal-block product.
constexpr int TS_THREADS = 1024;
constexpr int TS_G = TS_THREADS / CHOL_NB;  // 8 groups of 16
constexpr int TS_W = CHOL_NB / TS_G;        // 16

// Share of thread t in z = D v (D column-major NB x NB, D[m*NB + i] is
// element (i, m)): row i = t % NB, m in [16 (t / NB), +16).
struct TsDiag {
  double d[TS_W];
  __device__ __forceinline__ void load(const double *D, int nb, int tid) {
    const int i = tid % CHOL_NB, m0 = (tid / CHOL_NB) * TS_W;
#pragma unroll
    for (int m = 0; m < TS_W; ++m)
      d[m] = (i < nb && m0 + m < nb) ? D[(m0 + m) * CHOL_NB + i] : 0.0;
  }
  // z[i] = sum_m D[i][m] v[m]  (v, z, part in LDS; part: TS_G x NB)
  __device__ __forceinline__ void apply(const double *v, double *z, double (*part)[CHOL_NB],
                                        int tid) const {
    const int i = tid % CHOL_NB, g = tid / CHOL_NB, m0 = g * TS_W;
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < TS_W; ++m) s = fma(d[m], v[m0 + m], s);
    part[g][i] = s;
    __syncthreads();
    if (tid < CHOL_NB) {
      double t = 0.0;
#pragma unroll
      for (int q = 0; q < TS_G; ++q) t += part[q][tid];
      z[tid] = t;
    }
    __syncthreads();
  }
};

// out = D v for one block (the first block of the forward solve, the last of
// the backward one)
__global__ void __launch_bounds__(TS_THREADS)
    trsv_head_kernel(const double *D, const double *v, int nb, double *out, const double *skip) {
  if (skip && skip[0] != 0.0) return;
  __shared__ double vin[CHOL_NB], z[CHOL_NB], part[TS_G][CHOL_NB];
  const int tid = threadIdx.x;
  TsDiag ds;
  ds.load(D, nb, tid);
  if (tid < CHOL_NB) vin[tid] = tid < nb ? v[tid] : 0.0;
  __syncthreads();
  ds.apply(vin, z, part, tid);
  if (tid < nb) out[tid] = z[tid];
}

// Forward step k (y_k final in y):  b[r] -= sum_c L[r][k0+c] y_k[c] for rows r
// >= k1 = k0 + nbk.  Workgroup 0 takes block k+1 (rows [k1, k1+nb1)) and
// forms y_{k+1} = Dinv_{k+1} b_{k+1} from it; workgroups g >= 1 take rows
// k1 + nb1 + 128 (g - 1) + [0, 128).  Thread t: row t % 128, columns
// [16 (t / 128), +16) (lanes on consecutive rows: coalesced column reads).
__global__ void __launch_bounds__(TS_THREADS)
    trsv_fwd_kernel(const double *L, int64_t lda, int64_t n, int64_t k0, int nbk, int nb1,
                    const double *Dnext, double *b, double *y, const double *skip) {
  if (skip && skip[0] != 0.0) return;
  __shared__ double z[CHOL_NB], bn[CHOL_NB], part[TS_G][CHOL_NB];
  const int tid = threadIdx.x, i = tid % CHOL_NB, g = tid / CHOL_NB, c0 = g * TS_W;
  const bool lead = blockIdx.x == 0;
  const int64_t k1 = k0 + nbk;
  const int64_t r = lead ? k1 + i : k1 + nb1 + (int64_t)(blockIdx.x - 1) * CHOL_NB + i;
  const bool row_ok = lead ? (i < nb1) : (r < n);
  double lv[TS_W];
#pragma unroll
  for (int c = 0; c < TS_W; ++c)
    lv[c] = (row_ok && c0 + c < nbk) ? L[(k0 + c0 + c) * lda + r] : 0.0;
  TsDiag ds;
  if (lead) ds.load(Dnext, nb1, tid);
  double bi = 0.0;
  if (g == 0 && row_ok) bi = b[r];
  if (tid < CHOL_NB) z[tid] = tid < nbk ? y[k0 + tid] : 0.0;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < TS_W; ++c) s = fma(lv[c], z[c0 + c], s);
  part[g][i] = s;
  __syncthreads();
  double t = 0.0;
  if (g == 0) {
#pragma unroll
    for (int q = 0; q < TS_G; ++q) t += part[q][i];
  }
  if (!lead) {
    if (g == 0 && row_ok) b[r] = bi - t;
    return;
  }
  if (g == 0) bn[i] = row_ok ? bi - t : 0.0;
  __syncthreads();
  ds.apply(bn, z, part, tid);
  if (tid < nb1) y[k1 + tid] = z[tid];
}

// Backward step k (x_k final in x):  y[c] -= sum_r L[k0+r][c] x_k[r] for
// columns c < k0.  Workgroup 0 takes block k-1 (columns [k0-128, k0)) and
// forms x_{k-1} = Dinv_{k-1}^T y_{k-1} from it; workgroups g >= 1 take
// columns 128 (g - 1) + [0, 128).  Thread t: column t / 8, rows t % 8 + 8 j
// (per load, eight consecutive lanes read 64 contiguous bytes of a column).
__global__ void __launch_bounds__(TS_THREADS)
    trsv_bwd_kernel(const double *L, int64_t lda, int64_t k0, int nbk, const double *DTprev,
                    double *yv, double *x, const double *skip) {
  if (skip && skip[0] != 0.0) return;
  __shared__ double z[CHOL_NB], yn[CHOL_NB], part[TS_G][CHOL_NB];
  const int tid = threadIdx.x, cl = tid / TS_G, rg = tid % TS_G;
  const bool lead = blockIdx.x == 0;
  const int64_t c = lead ? k0 - CHOL_NB + cl : (int64_t)(blockIdx.x - 1) * CHOL_NB + cl;
  const bool col_ok = lead ? true : (c < k0 - CHOL_NB);
  double lv[TS_W];
  const double *col = L + c * lda + k0 + rg;
#pragma unroll
  for (int j = 0; j < TS_W; ++j) lv[j] = (col_ok && rg + TS_G * j < nbk) ? col[TS_G * j] : 0.0;
  TsDiag ds;
  if (lead) ds.load(DTprev, CHOL_NB, tid);
  double yc = 0.0;
  if (rg == 0 && col_ok) yc = yv[c];
  if (tid < CHOL_NB) z[tid] = tid < nbk ? x[k0 + tid] : 0.0;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < TS_W; ++j) s = fma(lv[j], z[rg + TS_G * j], s);
  // the eight row groups of a column are eight consecutive lanes
#pragma unroll
  for (int o = TS_G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (!lead) {
    if (rg == 0 && col_ok) yv[c] = yc - s;
    return;
  }
  if (rg == 0) yn[cl] = yc - s;
  __syncthreads();
  ds.apply(yn, z, part, tid);
  if (tid < CHOL_NB) x[k0 - CHOL_NB + tid] = z[tid];
}

__global__ void zero_i32(int32_t *p) { p[0] = 0; }

// [Dinv blocks | DinvT blocks | W: n x n column-major, the strictly lower
//  off-diagonal blocks of L (W[c*n + r] = L[r][c] for r in a later block)]
int64_t chol_dinv_bytes(int64_t n) {
  const int64_t nblk = (n + CHOL_NB - 1) / CHOL_NB;
  return (2 * nblk * CHOL_NB * CHOL_NB + n * n) * (int64_t)sizeof(double);
}

const double *chol_panels(int64_t n, const double *dinv) {
  const int64_t nblk = (n + CHOL_NB - 1) / CHOL_NB;
  return dinv + 2 * nblk * CHOL_NB * CHOL_NB;
}

#define VMCX_BLAS2(call)                                                      \
  do {                                                                       \
    rocblas_status s_ = (call);                                              \
    if (s_ != rocblas_status_success) {                                      \
      ::vmcx::set_error(std::string(__func__) + ": " #call " -> " +          \
                        rocblas_status_to_string(s_));                       \
      return 3;                                                              \
    }                                                                        \
  } while (0)

// Side streams and event pool for the look-ahead, one set per (device,
// caller stream): two callers on different streams never share a side
// stream or an event, so their factorisations cannot interleave.
struct LookAhead {
  hipStream_t side = nullptr;  // trailing SYRKs
  hipStream_t aux = nullptr;   // forward solve along the factorisation
  int32_t *busy = nullptr;     // CU code of the running diagonal kernel, or 0
  std::vector<hipEvent_t> ev;
};

static LookAhead *lookahead_for(hipStream_t caller, int nev) {
  static std::mutex mu;
  static std::map<std::pair<int, hipStream_t>, LookAhead *> per_stream;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  LookAhead *&la = per_stream[std::make_pair(dev, caller)];
  if (la == nullptr) {
    LookAhead *n = new LookAhead();
    if (hipStreamCreateWithFlags(&n->side, hipStreamNonBlocking) != hipSuccess) return nullptr;
    if (hipStreamCreateWithFlags(&n->aux, hipStreamNonBlocking) != hipSuccess) return nullptr;
    if (hipMalloc((void **)&n->busy, 256) != hipSuccess) return nullptr;
    if (hipMemset(n->busy, 0, 256) != hipSuccess) return nullptr;
    la = n;
  }
  while ((int)la->ev.size() < nev) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    la->ev.push_back(e);
  }
  return la;
}

// Right-looking blocked Cholesky with look-ahead and grouped trailing updates.
// The factor's off-diagonal blocks go to W (a separate n x n buffer), so the
// panel product needs no copy back and G = 2 consecutive panels sit side by
// side as one n x 2NB operand:
//   main stream : diag_k (factor + inverse) -> W_k = A_k,rest Dinv_k^T (GEMM)
//                 -> next block column k+1 -= W_(E,k] W_(E,k]^T (GEMM over the
//                 panels no trailing update has applied to it yet)
//   side stream : after the last panel e of each group, one matrix-core SYRK
//                 of the trailing matrix beyond block column e + G with the
//                 group's G panels (K = G NB)
// A trailing update therefore has G main-stream steps to finish before the
// main stream needs the columns it writes, and A is swept once per G panels.
// Block column c gets panel j from the SYRK of j's group when that group
// ends at e <= c - 1 - G, otherwise from the main stream's GEMM at step c-1;
// before that GEMM the main stream waits for the SYRK of group end
// E(c-1) = the last e <= c - 1 - G (earlier ones precede it on the side
// stream).  Regions written concurrently are disjoint: the SYRK of e writes
// block columns >= e + 1 + G, the main stream up to then only columns <= e + G.
//
// fwd_b != NULL: the forward solve y = L^-1 b (b overwritten, y = fwd_y) runs
// on a third stream as the blocks of L appear: its step k needs W_k and
// Dinv_{k+1}, so it follows diag_{k+1} and stays off the critical path.
int chol_factor(void *handle, int64_t n, double *A, int64_t lda, double *dinv,
                int32_t *info, hipStream_t st, float *diag_ms, double *fwd_b, double *fwd_y) {
  constexpr int64_t G = 2;
  rocblas_handle h = (rocblas_handle)handle;
  const int64_t nblk = (n + CHOL_NB - 1) / CHOL_NB;
  double *Dinv = dinv, *DinvT = dinv + nblk * CHOL_NB * CHOL_NB;
  double *W = dinv + 2 * nblk * CHOL_NB * CHOL_NB;
  const int64_t ldw = n;
  LookAhead *la = lookahead_for(st, (int)(3 * nblk + 3));
  hipEvent_t *evf = la ? la->ev.data() + 2 * nblk + 2 : nullptr;  // after diag_k
  VMCX_CHECK_ARG(la != nullptr, "look-ahead stream / handle creation failed");
  hipLaunchKernelGGL(zero_i32, dim3(1), dim3(1), 0, st, info);
  VMCX_LAUNCHED();
  const double one = 1.0, mone = -1.0, zero = 0.0;
  // side stream starts after everything already queued on the main stream
  VMCX_HIP(hipEventRecord(la->ev[0], st));
  VMCX_HIP(hipStreamWaitEvent(la->side, la->ev[0], 0));
  if (fwd_b) VMCX_HIP(hipStreamWaitEvent(la->aux, la->ev[0], 0));
  int64_t last_syrk = -1;    // last group end with a SYRK queued
  int64_t waited = -1;       // last group end the main stream waited for
  // optional timing of every diagonal-block launch (events on the main stream)
  std::vector<hipEvent_t> tev;
  if (diag_ms) {
    tev.resize(2 * nblk);
    for (auto &e : tev) VMCX_HIP(hipEventCreate(&e));
  }
  int64_t ndiag = 0;
  for (int64_t kb = 0; kb < nblk; ++kb) {
    const int64_t k0 = kb * CHOL_NB;
    const int nbk = (int)((n - k0 < CHOL_NB) ? n - k0 : CHOL_NB);
    double *Dk = Dinv + kb * CHOL_NB * CHOL_NB;
    if (diag_ms) VMCX_HIP(hipEventRecord(tev[2 * kb], st));
    hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(DIAG_THREADS), 0, st, A, lda, k0, nbk,
                       info, Dk, la->busy);
    VMCX_LAUNCHED();
    if (diag_ms) VMCX_HIP(hipEventRecord(tev[2 * kb + 1], st));
    ++ndiag;
    if (fwd_b) {  // forward-solve step kb - 1 (or the head) on the third stream
      VMCX_HIP(hipEventRecord(evf[kb], st));
      VMCX_HIP(hipStreamWaitEvent(la->aux, evf[kb], 0));
      if (kb == 0) {
        hipLaunchKernelGGL(trsv_head_kernel, dim3(1), dim3(TS_THREADS), 0, la->aux, Dinv, fwd_b,
                           nbk, fwd_y, nullptr);
      } else {
        const int64_t kp0 = k0 - CHOL_NB;
        const int64_t rest_p = n - k0 - nbk;
        hipLaunchKernelGGL(trsv_fwd_kernel, dim3((unsigned)(1 + (rest_p + CHOL_NB - 1) / CHOL_NB)),
                           dim3(TS_THREADS), 0, la->aux, W, ldw, n, kp0, CHOL_NB, nbk, Dk, fwd_b,
                           fwd_y, nullptr);
      }
      VMCX_LAUNCHED();
    }
    const int64_t k1 = k0 + nbk;
    const int64_t rest = n - k1;
    if (rest <= 0) break;
    // W_k = P L_kk^-T  (Dk = L_kk^-1, column-major NB x NB, zero above the diagonal)
    VMCX_BLAS2(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose,
                             (rocblas_int)rest, nbk, nbk, &one, A + k0 * lda + k1,
                             (rocblas_int)lda, Dk, CHOL_NB, &zero, W + k0 * ldw + k1,
                             (rocblas_int)ldw));
    // next block column k+1: the panels after the last group end E <= kb - G
    const int nb1 = (int)((rest < CHOL_NB) ? rest : CHOL_NB);
    const int64_t E = (kb - 2 * G + 1 >= 0) ? ((kb - 2 * G + 1) / G) * G + G - 1 : -1;
    if (E >= 0 && E > waited) {
      VMCX_HIP(hipStreamWaitEvent(st, la->ev[2 + 2 * E], 0));
      waited = E;
    }
    const int64_t p0 = (E + 1) * CHOL_NB;
    VMCX_BLAS2(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose,
                             (rocblas_int)rest, nb1, (rocblas_int)(k1 - p0), &mone,
                             W + p0 * ldw + k1, (rocblas_int)ldw, W + p0 * ldw + k1,
                             (rocblas_int)ldw, &one, A + k1 * lda + k1, (rocblas_int)lda));
    // group end: the group's trailing update beyond block column kb + G
    if ((kb + 1) % G == 0) {
      const int64_t s0 = (kb + 1 + G) * CHOL_NB;
      if (s0 < n) {
        VMCX_HIP(hipEventRecord(la->ev[1 + 2 * kb], st));
        VMCX_HIP(hipStreamWaitEvent(la->side, la->ev[1 + 2 * kb], 0));
        const int rc = syrk_f64(n - s0, G * CHOL_NB, W + (kb + 1 - G) * CHOL_NB * ldw + s0, ldw,
                                0, -1.0, 1.0, A + s0 * lda + s0, lda, 1, 0, la->side,
                                nullptr, la->busy);
        if (rc) return rc;
        VMCX_HIP(hipEventRecord(la->ev[2 + 2 * kb], la->side));
        last_syrk = kb;
      }
    }
  }
  if (last_syrk > waited) VMCX_HIP(hipStreamWaitEvent(st, la->ev[2 + 2 * last_syrk], 0));
  if (fwd_b) {
    VMCX_HIP(hipEventRecord(evf[nblk], la->aux));
    VMCX_HIP(hipStreamWaitEvent(st, evf[nblk], 0));
  }
  hipLaunchKernelGGL(chol_dinv_transpose_kernel, dim3((unsigned)(nblk * 64)), dim3(256), 0, st,
                     nblk, Dinv, DinvT);
  VMCX_LAUNCHED();
  if (diag_ms) {
    VMCX_HIP(hipEventSynchronize(tev[2 * ndiag - 1]));
    double tot = 0.0;
    for (int64_t k = 0; k < ndiag; ++k) {
      float ms = 0.0f;
      VMCX_HIP(hipEventElapsedTime(&ms, tev[2 * k], tev[2 * k + 1]));
      tot += ms;
    }
    for (auto &e : tev) (void)hipEventDestroy(e);
    *diag_ms = (float)(tot / (double)ndiag);
  }
  return 0;
}

// x = (L L^T)^-1 b ; tmp: n doubles; b is preserved
// Forward half: y = L^-1 b (b is overwritten).
int chol_solve_fwd(int64_t n, const double *L, int64_t lda, const double *dinv, double *b,
                   double *y, hipStream_t st, const double *skip) {
  const int64_t nblk = (n + CHOL_NB - 1) / CHOL_NB;
  const double *Dinv = dinv;
  const int64_t blk = (int64_t)CHOL_NB * CHOL_NB;
  auto nb_of = [&](int64_t kb) { return (int)((n - kb * CHOL_NB < CHOL_NB) ? n - kb * CHOL_NB : CHOL_NB); };
  hipLaunchKernelGGL(trsv_head_kernel, dim3(1), dim3(TS_THREADS), 0, st, Dinv, b, nb_of(0), y,
                     skip);
  VMCX_LAUNCHED();
  for (int64_t kb = 0; kb + 1 < nblk; ++kb) {
    const int64_t k0 = kb * CHOL_NB;
    const int nbk = nb_of(kb), nb1 = nb_of(kb + 1);
    const int64_t rest = n - k0 - nbk - nb1;
    const unsigned grid = (unsigned)(1 + (rest + CHOL_NB - 1) / CHOL_NB);
    hipLaunchKernelGGL(trsv_fwd_kernel, dim3(grid), dim3(TS_THREADS), 0, st, L, lda, n, k0, nbk,
                       nb1, Dinv + (kb + 1) * blk, b, y, skip);
    VMCX_LAUNCHED();
  }
  return 0;
}

// Backward half: x = L^-T y (y is overwritten).
int chol_solve_bwd(int64_t n, const double *L, int64_t lda, const double *dinv, double *y,
                   double *x, hipStream_t st, const double *skip) {
  const int64_t nblk = (n + CHOL_NB - 1) / CHOL_NB;
  const double *DinvT = dinv + nblk * CHOL_NB * CHOL_NB;
  const int64_t blk = (int64_t)CHOL_NB * CHOL_NB;
  const int64_t kl = nblk - 1;
  const int nbl = (int)(n - kl * CHOL_NB);
  hipLaunchKernelGGL(trsv_head_kernel, dim3(1), dim3(TS_THREADS), 0, st, DinvT + kl * blk,
                     y + kl * CHOL_NB, nbl, x + kl * CHOL_NB, skip);
  VMCX_LAUNCHED();
  for (int64_t kb = kl; kb >= 1; --kb) {
    const int64_t k0 = kb * CHOL_NB;
    const int nbk = (int)((n - k0 < CHOL_NB) ? n - k0 : CHOL_NB);
    const unsigned grid = (unsigned)(1 + (k0 - CHOL_NB + CHOL_NB - 1) / CHOL_NB);
    hipLaunchKernelGGL(trsv_bwd_kernel, dim3(grid), dim3(TS_THREADS), 0, st, L, lda, k0, nbk,
                       DinvT + (kb - 1) * blk, y, x, skip);
    VMCX_LAUNCHED();
  }
  return 0;
}

int chol_solve(int64_t n, const double *L, int64_t lda, const double *dinv,
               const double *b, double *x, double *tmp, double *tmp2, hipStream_t st,
               const double *skip) {
  VMCX_HIP(hipMemcpyAsync(tmp, b, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  int rc = chol_solve_fwd(n, L, lda, dinv, tmp, tmp2, st, skip);
  if (rc) return rc;
  return chol_solve_bwd(n, L, lda, dinv, tmp2, x, st, skip);
}

}  // namespace vmcx
