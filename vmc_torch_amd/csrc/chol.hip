// Blocked FP64 Cholesky factorisation and triangular solves for the SR
// system (column-major n x n, lower triangle).
//
// Factorisation, right-looking with block NB = 128:
//   for each diagonal block k:
//     chol_diag_kernel  : L_kk = chol(A_kk) in LDS (one workgroup), and
//                         Dinv_k = L_kk^-1 (kept for the solves)
//     rocblas_dtrsm     : A_ik <- A_ik L_kk^-T          (panel below)
//     rocblas_dsyrk     : A_22 <- A_22 - P P^T          (trailing update)
// The rocSOLVER panel kernel is latency bound at this size; the diagonal
// block here stays in LDS for all NB steps.
//
// Solves use the inverted diagonal blocks so that each block step is one
// launch: forward  y_k = Dinv_k (b_k) then b_rest -= L_rest,k y_k;
//         backward x_k = Dinv_k^T (y_k) then y_head -= L_k,head^T x_k.
#include <rocblas/rocblas.h>

#include <mutex>
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
//   leaf     : one wave factorises the 16 x 16 diagonal tile in registers
//              (lane i holds row i; columns broadcast with DPP row_newbcast)
//              and inverts it there; L_tt stays in the lower triangle of the
//              tile, the strictly lower part of X_tt = L_tt^-1 goes
//              transposed into its (unused) upper triangle, and the
//              reciprocal diagonal 1 / L_ii (= X_ii) into rdiag
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

// 1/d and 1/sqrt(d) for d > 0: hardware estimate and two Newton steps
__device__ __forceinline__ double rcp_f64(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = fma(-d, y, 1.0);
  y = fma(y, e, y);
  e = fma(-d, y, 1.0);
  return fma(y, e, y);
}

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

// value of v in lane S of each row of 16 lanes (DPP row_newbcast)
template <int S>
__device__ __forceinline__ double bcast16_c(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, 0x150 + S, 0xF, 0xF, false);
  const int hi =
      __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), 0x150 + S, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// s must fold to a constant (fully unrolled loops)
__device__ __forceinline__ double bcast16(double v, int s) {
  switch (s) {
    case 0: return bcast16_c<0>(v);
    case 1: return bcast16_c<1>(v);
    case 2: return bcast16_c<2>(v);
    case 3: return bcast16_c<3>(v);
    case 4: return bcast16_c<4>(v);
    case 5: return bcast16_c<5>(v);
    case 6: return bcast16_c<6>(v);
    case 7: return bcast16_c<7>(v);
    case 8: return bcast16_c<8>(v);
    case 9: return bcast16_c<9>(v);
    case 10: return bcast16_c<10>(v);
    case 11: return bcast16_c<11>(v);
    case 12: return bcast16_c<12>(v);
    case 13: return bcast16_c<13>(v);
    case 14: return bcast16_c<14>(v);
    default: return bcast16_c<15>(v);
  }
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Leaf of step kb (wave 0): factor and invert the tile at (kb, kb).
__device__ __forceinline__ void chol_leaf(double (*L)[CHOL_NB + 1], double *rdiag_all, int kb,
                                          int lane, int *bad) {
  const int i = lane & 15;
  const bool act = lane < LEAF;
  double r[LEAF];
#pragma unroll
  for (int c = 0; c < LEAF; ++c) r[c] = (c <= i) ? L[kb + i][kb + c] : 0.0;
  double myrs = 1.0;
  int badj = 0;
#pragma unroll
  for (int j = 0; j < LEAF; ++j) {
    // unscaled column j first (L_ij L_lj = a_ij a_lj / a_jj), so that the
    // broadcasts do not wait for the square root
    const double djj = bcast16(r[j], j);
    double col[LEAF];
#pragma unroll
    for (int l = j + 1; l < LEAF; ++l) col[l] = bcast16(r[j], l);
    const bool ok = djj > 0.0;
    if (!ok && badj == 0) badj = j + 1;
    const double inv = ok ? rcp_f64(djj) : 1.0;
    const double rs = ok ? rsqrt_f64(djj) : 1.0;
    const double t = r[j] * inv;
#pragma unroll
    for (int l = j + 1; l < LEAF; ++l) r[l] = (i >= l) ? fma(-t, col[l], r[l]) : r[l];
    r[j] = (i == j) ? (ok ? djj * rs : 1.0) : ((i > j) ? r[j] * rs : r[j]);
    if (i == j) myrs = rs;
  }
  // X = L_tt^-1 by rows: X_ii = 1 / L_ii, X_ic = -X_ii sum_{c <= q < i} L_iq X_qc
  double acc[LEAF];
#pragma unroll
  for (int c = 0; c < LEAF; ++c) acc[c] = 0.0;
#pragma unroll
  for (int q = 0; q < LEAF - 1; ++q) {
#pragma unroll
    for (int c = 0; c <= q; ++c) {
      const double v = (c == q) ? myrs : -myrs * acc[c];  // X[q][c] in lane q
      const double xb = bcast16(v, q);
      acc[c] = (i > q) ? fma(r[q], xb, acc[c]) : acc[c];
    }
  }
  if (act) {
#pragma unroll
    for (int c = 0; c < LEAF; ++c) {
      if (c <= i) L[kb + i][kb + c] = r[c];
      if (c < i) L[kb + c][kb + i] = -myrs * acc[c];
    }
    rdiag_all[kb + i] = myrs;
    if (lane == 0 && badj && *bad == 0) *bad = kb + badj;
  }
}

// X_tt (lower, the leaf inverse of tile t at kb) as a matrix-core B operand:
// element [k][j] of X_tt^T (transpose = 0) or of X_tt (transpose = 1)
__device__ __forceinline__ double leaf_inv(const double (*L)[CHOL_NB + 1],
                                          const double *rdiag_all, int kb, int row, int col) {
  // X_tt[row][col]
  if (row == col) return rdiag_all[kb + row];
  return (row > col) ? L[kb + col][kb + row] : 0.0;
}

__global__ void __launch_bounds__(DIAG_THREADS)
    chol_diag_kernel(double *A, int64_t lda, int64_t k0, int nbk, int32_t *info,
                     double *Dinv, double *DinvT) {
  __shared__ double L[CHOL_NB][CHOL_NB + 1];  // L[i][j], row i, column j
  __shared__ double W[CHOL_NB][LEAF + 1];     // inverse phase: W_T tiles
  __shared__ double rdiag_all[CHOL_NB];       // 1 / L[i][i]
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
      L[i][j] = v[q];
    }
  }
  if (tid == 0) bad = 0;
  __syncthreads();
  if (wave == 0) chol_leaf(L, rdiag_all, 0, lane, &bad);
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
        const double a = L[r0 + li][kb + kk];
        const double b = leaf_inv(L, rdiag_all, kb, li, kk);  // X_tt^T[kk][li]
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      }
      wave_lds_sync();
#pragma unroll
      for (int r = 0; r < 4; ++r) L[r0 + 4 * r + lk][kb + li] = acc[r];
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
        for (int r = 0; r < 4; ++r) acc[r] = L[r0 + 4 * r + lk][c0 + li];
#pragma unroll
        for (int s4 = 0; s4 < LEAF; s4 += 4) {
          const double a = -L[r0 + li][kb + s4 + lk];
          const double b = L[c0 + li][kb + s4 + lk];
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) L[r0 + 4 * r + lk][c0 + li] = acc[r];
      }
      if (wave == 0) {
        wave_lds_sync();
        chol_leaf(L, rdiag_all, kb + LEAF, lane, &bad);
      }
    }
    __syncthreads();
  }
  if (bad) {
    if (tid == 0 && info && info[0] == 0 && bad <= nbk) info[0] = (int32_t)(k0 + bad);
    return;
  }
  for (int idx = tid; idx < nbk * nbk; idx += DIAG_THREADS) {
    const int j = idx / nbk, i = idx - j * nbk;
    if (i >= j) Akk[(int64_t)j * lda + i] = L[i][j];
  }
  __syncthreads();
  // ---- X = L^-1 in place (off-diagonal tiles; the diagonal tiles already
  //      hold their inverses), tile column jt from the right
  for (int jt = NT - 2; jt >= 0; --jt) {
    const int jb = jt * LEAF, nt = NT - 1 - jt;
    if (wave < nt) {
      const int T = jt + 1 + wave;
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
      for (int K = jt + 1; K <= T; ++K) {
#pragma unroll
        for (int s4 = 0; s4 < LEAF; s4 += 4) {
          const int kk = s4 + lk;
          const double a = (K == T) ? leaf_inv(L, rdiag_all, LEAF * T, li, kk)
                                    : L[LEAF * T + li][LEAF * K + kk];
          const double b = L[LEAF * K + kk][jb + li];
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) W[LEAF * wave + 4 * r + lk][li] = acc[r];
    }
    __syncthreads();
    if (wave < nt) {
      const int T = jt + 1 + wave;
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s4 = 0; s4 < LEAF; s4 += 4) {
        const int kk = s4 + lk;
        const double a = -W[LEAF * wave + li][kk];
        const double b = leaf_inv(L, rdiag_all, jb, kk, li);  // X_jj[kk][li]
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) L[LEAF * T + 4 * r + lk][jb + li] = acc[r];
    }
    __syncthreads();
  }
  // Dinv[a*NB + b] = X[b][a];  DinvT[a*NB + b] = X[a][b]
  for (int idx = tid; idx < CHOL_NB * CHOL_NB; idx += DIAG_THREADS) {
    const int a = idx / CHOL_NB, b = idx - a * CHOL_NB;
    double xba = 0.0, xab = 0.0;
    if (a < nbk && b < nbk) {
      if ((a >> 4) == (b >> 4)) {
        const int tb = a & ~15;
        xba = leaf_inv(L, rdiag_all, tb, b - tb, a - tb);
        xab = leaf_inv(L, rdiag_all, tb, a - tb, b - tb);
      } else if (b > a) {
        xba = L[b][a];
      } else {
        xab = L[a][b];
      }
    }
    Dinv[idx] = xba;
    DinvT[idx] = xab;
  }
}

// forward block step: y_k = Dinv_k b_k ; b[r] -= sum_c L[r][k0+c] y_k[c] for r >= k0+nbk
__global__ void __launch_bounds__(256)
    trsv_fwd_kernel(const double *L, int64_t lda, int64_t n, int64_t k0, int nbk,
                    const double *Dinv, double *b, double *y, const double *skip) {
  if (skip && skip[0] != 0.0) return;
  __shared__ double z[CHOL_NB];
  const int tid = threadIdx.x;
  if (tid < CHOL_NB) {
    double s = 0.0;
    if (tid < nbk)
#pragma unroll 16
      for (int m = 0; m < nbk; ++m) s += Dinv[m * CHOL_NB + tid] * b[k0 + m];
    z[tid] = s;
  }
  __syncthreads();
  if (blockIdx.x == 0 && tid < nbk) y[k0 + tid] = z[tid];
  const int64_t r = k0 + nbk + (int64_t)blockIdx.x * blockDim.x + tid;
  if (r < n) {
    double s = 0.0;
#pragma unroll 16
    for (int c = 0; c < nbk; ++c) s += L[(k0 + c) * lda + r] * z[c];
    b[r] -= s;
  }
}

// backward block step: x_k = Dinv_k^T y_k ; y[c'] -= sum_r L[k0+r][c'] x_k[r] for c' < k0
// one wave per output column c'
__global__ void __launch_bounds__(256)
    trsv_bwd_kernel(const double *L, int64_t lda, int64_t k0, int nbk,
                    const double *DinvT, double *yv, double *x, const double *skip) {
  if (skip && skip[0] != 0.0) return;
  __shared__ double z[CHOL_NB];
  const int tid = threadIdx.x;
  if (tid < CHOL_NB) {
    double s = 0.0;
    if (tid < nbk)
#pragma unroll 16
      for (int m = 0; m < nbk; ++m) s += DinvT[m * CHOL_NB + tid] * yv[k0 + m];
    z[tid] = s;  // z[r] = sum_m (L^-1)[m][r] y[m]
  }
  __syncthreads();
  if (blockIdx.x == 0 && tid < nbk) x[k0 + tid] = z[tid];
  const int lane = tid & 63, wv = tid >> 6;
  const int64_t cp = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv;
  if (cp < k0) {
    const double *col = L + cp * lda + k0;
    double s = 0.0;
    for (int r = lane; r < nbk; r += 64) s += col[r] * z[r];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) yv[cp] -= s;
  }
}

__global__ void zero_i32(int32_t *p) { p[0] = 0; }

// [Dinv blocks | DinvT blocks | two n x NB panel buffers]
int64_t chol_dinv_bytes(int64_t n) {
  const int64_t nblk = (n + CHOL_NB - 1) / CHOL_NB;
  return (2 * nblk * CHOL_NB * CHOL_NB + 2 * n * CHOL_NB) * (int64_t)sizeof(double);
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

// Side stream, rocBLAS handle and event pool for the look-ahead (per device).
struct LookAhead {
  hipStream_t side = nullptr;
  rocblas_handle h2 = nullptr;
  std::vector<hipEvent_t> ev;
};

static LookAhead *lookahead_for_device(int nev) {
  static std::mutex mu;
  static std::unordered_map<int, LookAhead *> per_dev;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  LookAhead *la = per_dev[dev];
  if (la == nullptr) {
    la = new LookAhead();
    if (hipStreamCreateWithFlags(&la->side, hipStreamNonBlocking) != hipSuccess) return nullptr;
    if (rocblas_create_handle(&la->h2) != rocblas_status_success) return nullptr;
    rocblas_set_stream(la->h2, la->side);
    rocblas_set_pointer_mode(la->h2, rocblas_pointer_mode_host);
    per_dev[dev] = la;
  }
  while ((int)la->ev.size() < nev) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    la->ev.push_back(e);
  }
  return la;
}

// Right-looking blocked Cholesky with one-block look-ahead:
//   main stream : diag_k (factor + inverse) -> Pw = A_k,rest Dinv_k^T (one GEMM)
//                 -> [wait side stream of step k-1] -> next-column GEMM_k
//   side stream : [wait Pw] -> SYRK of the trailing matrix minus the next block
//                 column -> copy Pw into the factor
// so the large trailing updates overlap the latency-bound diagonal
// factorisation and panel of the next block.  The panel buffer alternates
// between two copies: step k+1 writes the one the side stream of step k-1
// read, which the main stream has waited for at step k.
int chol_factor(void *handle, int64_t n, double *A, int64_t lda, double *dinv,
                int32_t *info, hipStream_t st, float *diag_ms) {
  rocblas_handle h = (rocblas_handle)handle;
  const int64_t nblk = (n + CHOL_NB - 1) / CHOL_NB;
  double *Dinv = dinv, *DinvT = dinv + nblk * CHOL_NB * CHOL_NB;
  double *Pbuf[2] = {dinv + 2 * nblk * CHOL_NB * CHOL_NB,
                     dinv + 2 * nblk * CHOL_NB * CHOL_NB + n * CHOL_NB};
  const int ldp = (int)n;
  LookAhead *la = lookahead_for_device((int)(2 * nblk + 2));
  VMCX_CHECK_ARG(la != nullptr, "look-ahead stream / handle creation failed");
  hipLaunchKernelGGL(zero_i32, dim3(1), dim3(1), 0, st, info);
  VMCX_LAUNCHED();
  const double one = 1.0, mone = -1.0, zero = 0.0;
  // side stream starts after everything already queued on the main stream
  VMCX_HIP(hipEventRecord(la->ev[0], st));
  VMCX_HIP(hipStreamWaitEvent(la->side, la->ev[0], 0));
  hipEvent_t side_done = nullptr;  // side-stream work of the previous step
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
                       info, Dk, DinvT + kb * CHOL_NB * CHOL_NB);
    VMCX_LAUNCHED();
    if (diag_ms) VMCX_HIP(hipEventRecord(tev[2 * kb + 1], st));
    ++ndiag;
    const int64_t rest = n - k0 - nbk;
    if (rest <= 0) break;
    double *P = A + k0 * lda + k0 + nbk;
    double *Pw = Pbuf[kb & 1];
    // Pw = P L_kk^-T  (Dk = L_kk^-1, column-major NB x NB, zero above the diagonal)
    VMCX_BLAS2(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose,
                             (rocblas_int)rest, nbk, nbk, &one, P, (rocblas_int)lda, Dk,
                             CHOL_NB, &zero, Pw, ldp));
    hipEvent_t panel_done = la->ev[1 + 2 * kb];
    VMCX_HIP(hipEventRecord(panel_done, st));
    // next block column: A[k1:, k1:k1+nb1] -= Pw Pw[0:nb1]^T (main stream)
    const int64_t k1 = k0 + nbk;
    const int nb1 = (int)((n - k1 < CHOL_NB) ? n - k1 : CHOL_NB);
    if (side_done) VMCX_HIP(hipStreamWaitEvent(st, side_done, 0));
    VMCX_BLAS2(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose,
                             (rocblas_int)rest, nb1, nbk, &mone, Pw, ldp, Pw, ldp, &one,
                             A + k1 * lda + k1, (rocblas_int)lda));
    // side stream: the rest of the trailing matrix, then the panel into the factor
    VMCX_HIP(hipStreamWaitEvent(la->side, panel_done, 0));
    const int64_t rest2 = rest - nb1;
    if (rest2 > 0) {
      double *A22 = A + (k1 + nb1) * lda + k1 + nb1;
      VMCX_BLAS2(rocblas_dsyrk(la->h2, rocblas_fill_lower, rocblas_operation_none,
                               (rocblas_int)rest2, nbk, &mone, Pw + nb1, ldp, &one, A22,
                               (rocblas_int)lda));
    }
    VMCX_HIP(hipMemcpy2DAsync(P, (size_t)lda * sizeof(double), Pw, (size_t)ldp * sizeof(double),
                              (size_t)rest * sizeof(double), (size_t)nbk,
                              hipMemcpyDeviceToDevice, la->side));
    side_done = la->ev[2 + 2 * kb];
    VMCX_HIP(hipEventRecord(side_done, la->side));
  }
  if (side_done) VMCX_HIP(hipStreamWaitEvent(st, side_done, 0));
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
int chol_solve(int64_t n, const double *L, int64_t lda, const double *dinv,
               const double *b, double *x, double *tmp, double *tmp2, hipStream_t st,
               const double *skip) {
  const int64_t nblk = (n + CHOL_NB - 1) / CHOL_NB;
  const double *Dinv = dinv, *DinvT = dinv + nblk * CHOL_NB * CHOL_NB;
  VMCX_HIP(hipMemcpyAsync(tmp, b, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  // forward: tmp2 = L^-1 b
  for (int64_t kb = 0; kb < nblk; ++kb) {
    const int64_t k0 = kb * CHOL_NB;
    const int nbk = (int)((n - k0 < CHOL_NB) ? n - k0 : CHOL_NB);
    const int64_t rest = n - k0 - nbk;
    const unsigned grid = (unsigned)((rest + 255) / 256 > 0 ? (rest + 255) / 256 : 1);
    hipLaunchKernelGGL(trsv_fwd_kernel, dim3(grid), dim3(256), 0, st, L, lda, n, k0, nbk,
                       Dinv + kb * CHOL_NB * CHOL_NB, tmp, tmp2, skip);
    VMCX_LAUNCHED();
  }
  // backward: x = L^-T tmp2
  for (int64_t kb = nblk - 1; kb >= 0; --kb) {
    const int64_t k0 = kb * CHOL_NB;
    const int nbk = (int)((n - k0 < CHOL_NB) ? n - k0 : CHOL_NB);
    const unsigned grid = (unsigned)((k0 + 3) / 4 > 0 ? (k0 + 3) / 4 : 1);
    hipLaunchKernelGGL(trsv_bwd_kernel, dim3(grid), dim3(256), 0, st, L, lda, k0, nbk,
                       DinvT + kb * CHOL_NB * CHOL_NB, tmp2, x, skip);
    VMCX_LAUNCHED();
  }
  return 0;
}

}  // namespace vmcx
