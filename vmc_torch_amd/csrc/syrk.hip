// FP64 symmetric rank-K update on the matrix cores:
//   C = beta C + alpha M M^T   (lower triangle of the n x n column-major C)
// with M (n x K) stored either row-major (M(i, k) = X[i*ldx + k]: the
// sample-major O of the SR Gram) or column-major (M(i, k) = X[k*ldx + i]:
// the panels of the blocked Cholesky).
//
// One workgroup (8 waves, two per SIMD: one wave per SIMD leaves the FP64
// matrix pipe about half idle) per 128 x 128 lower tile, one launch for the
// whole triangle; the reduction is staged through LDS in chunks of 16
// (double buffered, one barrier per chunk).  Wave w owns the 64 x 32 block
// (rows 64 (w >> 2), columns 32 (w & 3)) as 4 x 2 blocks of
// v_mfma_f64_16x16x4f64, accumulators in VGPRs; on a diagonal tile the two
// strictly upper blocks are skipped and the straddling ones are masked on
// the store.
//
// Operand roles: the A operand carries the tile's column block and the B
// operand its row block, so that D[4 r + l / 16][l % 16] of lane l is
// C[row l % 16][col 4 r + l / 16]: sixteen consecutive lanes store sixteen
// consecutive rows of one column.
//
// Splits (gridDim.y > 1): split s reduces k in [s kc, min(K, (s+1) kc)) and
// writes its partial to C (s = 0) or Crest + (s - 1) cstride (alpha, beta
// apply per split).
#include "kernels.h"

namespace vmcx {

constexpr int SK_T = 128;       // output tile
constexpr int SK_KC = 16;       // reduction chunk in LDS
// LDS chunk layouts (doubles), both SK_T x SK_KC = 2304 slots with padding:
//   column-major M: [k][i], row 144 (k rows 32 banks apart: conflict-free
//                   fragment reads, the global rows arrive i-contiguous)
//   row-major M:    [i][k], row 18 (fragment reads of 16 rows x 4 k hit 64
//                   distinct banks; the global rows arrive k-contiguous, so
//                   the [k][i] layout would make the stores 16-way conflicts)
constexpr int SK_LDK = 144;
constexpr int SK_LDI = 18;
constexpr int SK_SLOTS = SK_KC * SK_LDK;
static_assert(SK_T * SK_LDI == SK_SLOTS, "both layouts fill the same LDS slots");
template <bool ROWMAJ>
__device__ __forceinline__ int sk_idx(int i, int k) {
  return ROWMAJ ? i * SK_LDI + k : k * SK_LDK + i;
}
#ifndef SK_WAVES_C
#define SK_WAVES_C 4   // waves along the tile's columns (x 2 along its rows)
#endif
constexpr int SK_WC = SK_WAVES_C;
constexpr int SK_WR = 2;
constexpr int SK_THREADS = 64 * SK_WR * SK_WC;
constexpr int SK_BR = SK_T / SK_WR / 16;   // 16-row blocks per wave
constexpr int SK_BC = SK_T / SK_WC / 16;   // 16-column blocks per wave
// two workgroups per CU (LDS) -> SK_WC waves per SIMD
constexpr int SK_PER = SK_T * SK_KC / SK_THREADS;  // elements per thread per chunk

typedef double sk_dbl4 __attribute__((ext_vector_type(4)));

// Per-thread share of a 128 x 16 chunk of M (rows [r0, r0 + 128), k from
// kb + 16 c): SK_PER elements q at rows / k offsets fixed for the whole
// reduction, so only one pointer moves.
//   row-major:    k = kk (SK_KC consecutive k of a row), rows i0 + (512 / SK_KC) q
//   column-major: row i (128 consecutive rows per 2 waves), k = kk0 + 4 q
template <bool ROWMAJ>
struct SkStream {
  const double *p;   // element q = 0 of the current chunk
  int64_t qstep;     // pointer step between q
  int64_t cstep;     // pointer step between chunks
  int64_t kpos;      // k of element q = 0 of the current chunk
  int li, lk;        // LDS slot of element q = 0: S[lk][li]
  bool row_ok[SK_PER];

  __device__ __forceinline__ SkStream(const double *X, int64_t ldx, int64_t n, int64_t r0,
                                      int64_t kb, int tid) {
    if (ROWMAJ) {
      lk = tid % SK_KC;
      li = tid / SK_KC;
      p = X + (r0 + li) * ldx + kb + lk;
      qstep = (int64_t)(SK_THREADS / SK_KC) * ldx;
      cstep = SK_KC;
      kpos = kb + lk;
#pragma unroll
      for (int q = 0; q < SK_PER; ++q) row_ok[q] = r0 + li + q * (SK_THREADS / SK_KC) < n;
    } else {
      li = tid % SK_T;
      lk = tid / SK_T;
      p = X + (kb + lk) * ldx + r0 + li;
      qstep = (int64_t)(SK_THREADS / SK_T) * ldx;
      cstep = SK_KC * ldx;
      kpos = kb + lk;
#pragma unroll
      for (int q = 0; q < SK_PER; ++q) row_ok[q] = r0 + li < n;
    }
  }
  // load the current chunk (zero outside n x kend) and advance
  __device__ __forceinline__ void fetch(int64_t kend, double (&v)[SK_PER]) {
#pragma unroll
    for (int q = 0; q < SK_PER; ++q) {
      const int64_t k = ROWMAJ ? kpos : kpos + q * (SK_THREADS / SK_T);
      v[q] = (row_ok[q] && k < kend) ? p[q * qstep] : 0.0;
    }
    p += cstep;
    kpos += SK_KC;
  }
  __device__ __forceinline__ void stash(double *S, const double (&v)[SK_PER]) const {
#pragma unroll
    for (int q = 0; q < SK_PER; ++q) {
      if (ROWMAJ)
        S[sk_idx<true>(li + q * (SK_THREADS / SK_KC), lk)] = v[q];
      else
        S[sk_idx<false>(li, lk + q * (SK_THREADS / SK_T))] = v[q];
    }
  }
};

template <bool ROWMAJ>
__global__ void __launch_bounds__(SK_THREADS)
    __attribute__((amdgpu_waves_per_eu(SK_WC, SK_WC)))
    syrk_f64_kernel(int64_t n, int64_t K, const double *__restrict__ X, int64_t ldx,
                    int64_t kc, double alpha, double beta, double *C, int64_t ldc,
                    double *Crest, int64_t cstride, int ntiles, const int32_t *yield) {
  __shared__ double sR[2][SK_SLOTS];  // M rows of the tile's row block
  __shared__ double sC[2][SK_SLOTS];  // M rows of the tile's column block
  // XCD-aware order: consecutive workgroups go to different XCDs, so hand
  // each XCD a contiguous run of tiles (shared row / column blocks in its L2)
  const int per = (ntiles + 7) >> 3;
  const int tile = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (tile >= ntiles) return;
  int I = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= tile) ++I;
  while (I * (I + 1) / 2 > tile) --I;
  const int J = tile - I * (I + 1) / 2;
  const bool diag = (I == J);
  const int64_t r0 = (int64_t)I * SK_T, c0 = (int64_t)J * SK_T;
  const int64_t kb = (int64_t)blockIdx.y * kc;
  const int64_t ke = (kb + kc < K) ? kb + kc : K;
  if (blockIdx.y > 0) C = Crest + (int64_t)(blockIdx.y - 1) * cstride;

  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wr = w / SK_WC, wc = w % SK_WC;
  const int rw0 = 16 * SK_BR * wr, cw0 = 16 * SK_BC * wc;  // the wave's block in the tile
  const bool active = !(diag && cw0 >= rw0 + 16 * SK_BR);

  // |alpha| = 1 and beta != 0: the accumulators start from alpha beta C, read
  // here so that its latency overlaps the first chunk (the store is then
  // alpha acc = beta C + alpha M M^T); otherwise C is read in the epilogue
  const bool fold = beta != 0.0 && (alpha == 1.0 || alpha == -1.0);
  const double cscale = alpha * beta;
  // lane's view of C: element (a, b, r) at pc[(16 a + 4 r) ldc + 16 b], row
  // gi0 + 16 b, column gj0 + 16 a + 4 r
  const int64_t gi0 = r0 + rw0 + li, gj0 = c0 + cw0 + lk;
  double *pc = C + gj0 * ldc + gi0;
  sk_dbl4 acc[SK_BC][SK_BR];
#pragma unroll
  for (int a = 0; a < SK_BC; ++a)
#pragma unroll
    for (int b = 0; b < SK_BR; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = gi0 + 16 * b, gj = gj0 + 16 * a + 4 * r;
        acc[a][b][r] = (fold && active && gi < n && gj <= gi)
                           ? cscale * pc[(int64_t)(16 * a + 4 * r) * ldc + 16 * b]
                           : 0.0;
      }

  // yielding to a diagonal factorisation kernel on this CU: the flag is
  // loaded beside each chunk's fetch and tested at the next chunk, so the
  // test never waits on memory unless it sleeps
  const int32_t mycu = yield ? cu_code() : 0;
  int32_t yseen =
      yield ? __hip_atomic_load(yield, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
  const int64_t nck = (ke > kb) ? (ke - kb + SK_KC - 1) / SK_KC : 0;
  SkStream<ROWMAJ> gr(X, ldx, n, r0, kb, tid), gc(X, ldx, n, c0, kb, tid);
  double vr[SK_PER], vc[SK_PER];
  if (nck > 0) {
    gr.fetch(ke, vr);
    if (!diag) gc.fetch(ke, vc);
    gr.stash(sR[0], vr);
    if (!diag) gc.stash(sC[0], vc);
  }
  __syncthreads();
  for (int64_t c = 0; c < nck; ++c) {
    const int buf = (int)(c & 1);
    const bool more = c + 1 < nck;
    if (yield && yseen == mycu) {
      for (int spin = 0; spin < 48; ++spin) {   // 48 x ~3.4 us at most
        __builtin_amdgcn_s_sleep(127);
        if (__hip_atomic_load(yield, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != mycu) break;
      }
    }
    if (more) {
      gr.fetch(ke, vr);
      if (!diag) gc.fetch(ke, vc);
    }
    if (yield) yseen = __hip_atomic_load(yield, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (active) {
      const double *R = sR[buf];
      const double *Cm = diag ? sR[buf] : sC[buf];
#pragma unroll
      for (int s4 = 0; s4 < SK_KC; s4 += 4) {
        double fa[SK_BC], fb[SK_BR];
#pragma unroll
        for (int q = 0; q < SK_BC; ++q)  // A: column block
          fa[q] = Cm[sk_idx<ROWMAJ>(cw0 + 16 * q + li, s4 + lk)];
#pragma unroll
        for (int q = 0; q < SK_BR; ++q)  // B: row block
          fb[q] = R[sk_idx<ROWMAJ>(rw0 + 16 * q + li, s4 + lk)];
#pragma unroll
        for (int a = 0; a < SK_BC; ++a)
#pragma unroll
          for (int b = 0; b < SK_BR; ++b)
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[b], acc[a][b], 0, 0, 0);
      }
    }
    if (more) {
      gr.stash(sR[buf ^ 1], vr);
      if (!diag) gc.stash(sC[buf ^ 1], vc);
    }
    __syncthreads();
  }
  if (!active) return;
  // acc[a][b] lane l, reg r: C[row rw0 + 16 b + li][col cw0 + 16 a + 4 r + lk]
#pragma unroll
  for (int a = 0; a < SK_BC; ++a)
#pragma unroll
    for (int b = 0; b < SK_BR; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = gi0 + 16 * b, gj = gj0 + 16 * a + 4 * r;
        if (gi < n && gj <= gi) {
          double *p = pc + (int64_t)(16 * a + 4 * r) * ldc + 16 * b;
          const double v = alpha * acc[a][b][r];
          *p = (beta == 0.0 || fold) ? v : fma(beta, *p, v);
        }
      }
}

int syrk_f64(int64_t n, int64_t K, const double *X, int64_t ldx, int rowmajor, double alpha,
             double beta, double *C, int64_t ldc, int splits, int64_t cstride, hipStream_t st,
             double *Crest, const int32_t *yield) {
  VMCX_CHECK_ARG(n >= 0 && K >= 0, "n and K must be >= 0");
  if (n == 0) return 0;
  VMCX_CHECK_ARG(X != nullptr || K == 0, "X is NULL");
  VMCX_CHECK_ARG(C != nullptr, "C is NULL");
  VMCX_CHECK_ARG(ldc >= n, "ldc < n");
  VMCX_CHECK_ARG(rowmajor ? ldx >= K : ldx >= n, "ldx too small for the layout");
  VMCX_CHECK_ARG(splits >= 1 && splits <= 65535, "splits must be in [1, 65535]");
  VMCX_CHECK_ARG(splits == 1 || cstride >= ldc * n, "cstride smaller than one C");
  if (Crest == nullptr) Crest = C + cstride;
  const int64_t nt = (n + SK_T - 1) / SK_T;
  VMCX_CHECK_ARG(nt * (nt + 1) / 2 < (int64_t)1 << 30, "n too large for one launch");
  const int ntiles = (int)(nt * (nt + 1) / 2);
  const int64_t kc = (K + splits - 1) / splits;
  const unsigned gx = (unsigned)(((ntiles + 7) / 8) * 8);
  if (rowmajor)
    hipLaunchKernelGGL(syrk_f64_kernel<true>, dim3(gx, (unsigned)splits), dim3(SK_THREADS), 0, st,
                       n, K, X, ldx, kc, alpha, beta, C, ldc, Crest, cstride, ntiles, yield);
  else
    hipLaunchKernelGGL(syrk_f64_kernel<false>, dim3(gx, (unsigned)splits), dim3(SK_THREADS), 0,
                       st, n, K, X, ldx, kc, alpha, beta, C, ldc, Crest, cstride, ntiles,
                       yield);
  VMCX_LAUNCHED();
  return 0;
}

}  // namespace vmcx
