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
// The 128 x 128 diagonal block is factorised in LDS in 16-column steps:
//   leaf   : one wave factorises the 16 x 16 diagonal tile in registers
//            (lane i holds row i; columns broadcast with shuffles)
//   panel  : one thread per row solves x L_tile^T = a by substitution
//   update : rank-16 update of the trailing lower triangle
// and its inverse is formed in place column-block by column-block from the
// right (blocked LAPACK trtri order), with 16 x 16 leaf inverses.
constexpr int LEAF = 16;

__global__ void __launch_bounds__(1024)
    chol_diag_kernel(double *A, int64_t lda, int64_t k0, int nbk, int32_t *info) {
  __shared__ double L[CHOL_NB][CHOL_NB + 1];  // L[i][j], row i, column j
  __shared__ double rdiag[LEAF];              // 1 / L[kb+c][kb+c] of the current tile
  __shared__ int bad;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6;
  double *Akk = A + k0 * lda + k0;
  for (int idx = tid; idx < CHOL_NB * CHOL_NB; idx += nt) {
    const int j = idx / CHOL_NB, i = idx - j * CHOL_NB;
    double v = 0.0;
    if (i < nbk && j < nbk) v = (i >= j) ? Akk[(int64_t)j * lda + i] : 0.0;
    else if (i == j) v = 1.0;  // identity padding keeps the leaves regular
    L[i][j] = v;
  }
  if (tid == 0) bad = 0;
  __syncthreads();

  for (int kb = 0; kb < CHOL_NB; kb += LEAF) {
    // ---- leaf: rows kb..kb+15 by lanes 0..15 of wave 0
    if (wave == 0) {
      const int i = lane;  // lanes >= LEAF carry dummies
      double r[LEAF];
#pragma unroll
      for (int c = 0; c < LEAF; ++c)
        r[c] = (i < LEAF && c <= i) ? L[kb + i][kb + c] : 0.0;
#pragma unroll
      for (int j = 0; j < LEAF; ++j) {
        const double djj = __shfl(r[j], j, 64);
        const bool ok = djj > 0.0;
        if (!ok && lane == 0 && bad == 0) bad = kb + j + 1;
        const double sd = ok ? sqrt(djj) : 1.0;
        const double rs = 1.0 / sd;
        if (i == j) r[j] = sd;
        if (i > j && i < LEAF) r[j] *= rs;
        if (lane == j) rdiag[j] = rs;
#pragma unroll
        for (int l = j + 1; l < LEAF; ++l) {
          const double llj = __shfl(r[j], l, 64);
          if (i >= l && i < LEAF) r[l] -= r[j] * llj;
        }
      }
      if (i < LEAF) {
#pragma unroll
        for (int c = 0; c < LEAF; ++c)
          if (c <= i) L[kb + i][kb + c] = r[c];
      }
    }
    __syncthreads();
    // ---- panel: rows below the tile, x L_tile^T = a (one thread per row)
    for (int row = kb + LEAF + tid; row < CHOL_NB; row += nt) {
      double x[LEAF];
#pragma unroll
      for (int c = 0; c < LEAF; ++c) {
        double v = L[row][kb + c];
#pragma unroll
        for (int m = 0; m < c; ++m) v -= x[m] * L[kb + c][kb + m];
        x[c] = v * rdiag[c];
      }
#pragma unroll
      for (int c = 0; c < LEAF; ++c) L[row][kb + c] = x[c];
    }
    __syncthreads();
    // ---- trailing update: L[i][l] -= P_i . P_l for kb+16 <= l <= i,
    //      8 threads per row, columns strided by 8
    {
      const int base = kb + LEAF;
      const int i = base + (tid >> 3);
      if (i < CHOL_NB) {
        double pi[LEAF];
#pragma unroll
        for (int m = 0; m < LEAF; ++m) pi[m] = L[i][kb + m];
        for (int l = base + (tid & 7); l <= i; l += 8) {
          double d0 = 0.0, d1 = 0.0;
#pragma unroll
          for (int m = 0; m < LEAF; m += 2) {
            d0 += pi[m] * L[l][kb + m];
            d1 += pi[m + 1] * L[l][kb + m + 1];
          }
          L[i][l] -= d0 + d1;
        }
      }
    }
    __syncthreads();
  }
  if (bad) {
    if (tid == 0 && info && info[0] == 0 && bad <= nbk) info[0] = (int32_t)(k0 + bad);
    return;
  }
  for (int idx = tid; idx < nbk * nbk; idx += nt) {
    const int j = idx / nbk, i = idx - j * nbk;
    if (i >= j) Akk[(int64_t)j * lda + i] = L[i][j];
  }
}

// Inverse of every factorised diagonal block (one workgroup each, all in
// parallel after the factorisation), in place in LDS, column blocks from the
// right:  X_jj = inv(L_jj);  X_{below,j} = -X_{below,below} L_{below,j} X_jj
__global__ void __launch_bounds__(256)
    chol_diag_inverse_kernel(const double *A, int64_t lda, int64_t n, double *Dinv,
                             double *DinvT) {
  __shared__ double L[CHOL_NB][CHOL_NB + 1];
  __shared__ double W[CHOL_NB][LEAF + 1];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t kb = blockIdx.x, k0 = kb * CHOL_NB;
  const int nbk = (int)((n - k0 < CHOL_NB) ? n - k0 : CHOL_NB);
  const double *Akk = A + k0 * lda + k0;
  for (int idx = tid; idx < CHOL_NB * CHOL_NB; idx += nt) {
    const int j = idx / CHOL_NB, i = idx - j * CHOL_NB;
    double v = 0.0;
    if (i < nbk && j < nbk) v = (i >= j) ? Akk[(int64_t)j * lda + i] : 0.0;
    else if (i == j) v = 1.0;
    L[i][j] = v;
  }
  __syncthreads();
  for (int jb = CHOL_NB - LEAF; jb >= 0; jb -= LEAF) {
    const int base = jb + LEAF, rows = CHOL_NB - base;
    // W = X_{below,below} (already inverted, lower) * L_{below,j}
    for (int idx = tid; idx < rows * LEAF; idx += nt) {
      const int ii = idx / LEAF, c = idx - ii * LEAF;
      const int i = base + ii;
      double s = 0.0;
      for (int q = base; q <= i; ++q) s += L[i][q] * L[q][jb + c];
      W[ii][c] = s;
    }
    // leaf inverse of the 16 x 16 tile by 16 lanes (column c each)
    if (wave == 0) {
      const int c = lane;
      double x[LEAF];
#pragma unroll
      for (int i = 0; i < LEAF; ++i) {
        double v = (i == c) ? 1.0 : 0.0;
#pragma unroll
        for (int q = 0; q < i; ++q) v -= L[jb + i][jb + q] * x[q];
        x[i] = (c < LEAF && i >= c) ? v / L[jb + i][jb + i] : 0.0;
      }
      __builtin_amdgcn_wave_barrier();
      if (c < LEAF) {
#pragma unroll
        for (int i = 0; i < LEAF; ++i)
          if (i >= c) L[jb + i][jb + c] = x[i];
      }
    }
    __syncthreads();
    // X_{below,j} = -W * X_jj
    for (int idx = tid; idx < rows * LEAF; idx += nt) {
      const int ii = idx / LEAF, c = idx - ii * LEAF;
      double s = 0.0;
      for (int q = c; q < LEAF; ++q) s += W[ii][q] * L[jb + q][jb + c];
      L[base + ii][jb + c] = -s;
    }
    __syncthreads();
  }
  for (int idx = tid; idx < CHOL_NB * CHOL_NB; idx += nt) {
    const int a = idx / CHOL_NB, b = idx - a * CHOL_NB;
    // Dinv[m*NB + c] = X[c][m] (m = a, c = b); DinvT[c*NB + m] = X[c][m] (c = a, m = b)
    Dinv[kb * CHOL_NB * CHOL_NB + idx] = (a < nbk && b < nbk && b >= a) ? L[b][a] : 0.0;
    DinvT[kb * CHOL_NB * CHOL_NB + idx] = (a < nbk && b < nbk && a >= b) ? L[a][b] : 0.0;
  }
}


// forward block step: y_k = Dinv_k b_k ; b[r] -= sum_c L[r][k0+c] y_k[c] for r >= k0+nbk
__global__ void __launch_bounds__(256)
    trsv_fwd_kernel(const double *L, int64_t lda, int64_t n, int64_t k0, int nbk,
                    const double *Dinv, double *b, double *y) {
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
                    const double *DinvT, double *yv, double *x) {
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

int64_t chol_dinv_bytes(int64_t n) {
  const int64_t nblk = (n + CHOL_NB - 1) / CHOL_NB;
  return 2 * nblk * CHOL_NB * CHOL_NB * (int64_t)sizeof(double);
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
//   main stream : diag_k -> trsm_k -> [wait syrk_{k-1}] -> next-column GEMM_k
//   side stream : [wait trsm_k] -> SYRK of the trailing matrix minus the next
//                 block column
// so the large trailing updates overlap the latency-bound diagonal
// factorisation and panel solve of the next block.
int chol_factor(void *handle, int64_t n, double *A, int64_t lda, double *dinv,
                int32_t *info, hipStream_t st) {
  rocblas_handle h = (rocblas_handle)handle;
  const int64_t nblk = (n + CHOL_NB - 1) / CHOL_NB;
  double *Dinv = dinv, *DinvT = dinv + nblk * CHOL_NB * CHOL_NB;
  LookAhead *la = lookahead_for_device((int)(2 * nblk + 2));
  VMCX_CHECK_ARG(la != nullptr, "look-ahead stream / handle creation failed");
  hipLaunchKernelGGL(zero_i32, dim3(1), dim3(1), 0, st, info);
  VMCX_LAUNCHED();
  const double one = 1.0, mone = -1.0;
  // side stream starts after everything already queued on the main stream
  VMCX_HIP(hipEventRecord(la->ev[0], st));
  VMCX_HIP(hipStreamWaitEvent(la->side, la->ev[0], 0));
  hipEvent_t syrk_done = nullptr;  // SYRK of the previous step (side stream)
  for (int64_t kb = 0; kb < nblk; ++kb) {
    const int64_t k0 = kb * CHOL_NB;
    const int nbk = (int)((n - k0 < CHOL_NB) ? n - k0 : CHOL_NB);
    hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(1024), 0, st, A, lda, k0, nbk, info);
    VMCX_LAUNCHED();
    const int64_t rest = n - k0 - nbk;
    if (rest <= 0) break;
    double *Lkk = A + k0 * lda + k0;
    double *P = A + k0 * lda + k0 + nbk;
    VMCX_BLAS2(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower,
                             rocblas_operation_transpose, rocblas_diagonal_non_unit,
                             (rocblas_int)rest, nbk, &one, Lkk, (rocblas_int)lda, P,
                             (rocblas_int)lda));
    hipEvent_t panel_done = la->ev[1 + 2 * kb];
    VMCX_HIP(hipEventRecord(panel_done, st));
    // next block column: A[k1:, k1:k1+nb1] -= P P[0:nb1]^T (main stream)
    const int64_t k1 = k0 + nbk;
    const int nb1 = (int)((n - k1 < CHOL_NB) ? n - k1 : CHOL_NB);
    if (syrk_done) VMCX_HIP(hipStreamWaitEvent(st, syrk_done, 0));
    VMCX_BLAS2(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose,
                             (rocblas_int)rest, nb1, nbk, &mone, P, (rocblas_int)lda, P,
                             (rocblas_int)lda, &one, A + k1 * lda + k1, (rocblas_int)lda));
    // the rest of the trailing matrix on the side stream
    const int64_t rest2 = rest - nb1;
    if (rest2 > 0) {
      VMCX_HIP(hipStreamWaitEvent(la->side, panel_done, 0));
      double *P2 = P + nb1;
      double *A22 = A + (k1 + nb1) * lda + k1 + nb1;
      VMCX_BLAS2(rocblas_dsyrk(la->h2, rocblas_fill_lower, rocblas_operation_none,
                               (rocblas_int)rest2, nbk, &mone, P2, (rocblas_int)lda, &one,
                               A22, (rocblas_int)lda));
      syrk_done = la->ev[2 + 2 * kb];
      VMCX_HIP(hipEventRecord(syrk_done, la->side));
    } else {
      syrk_done = nullptr;
    }
  }
  if (syrk_done) VMCX_HIP(hipStreamWaitEvent(st, syrk_done, 0));
  hipLaunchKernelGGL(chol_diag_inverse_kernel, dim3((unsigned)nblk), dim3(256), 0, st, A, lda,
                     n, Dinv, DinvT);
  VMCX_LAUNCHED();
  return 0;
}

// x = (L L^T)^-1 b ; tmp: n doubles; b is preserved
int chol_solve(int64_t n, const double *L, int64_t lda, const double *dinv,
               const double *b, double *x, double *tmp, double *tmp2, hipStream_t st) {
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
                       Dinv + kb * CHOL_NB * CHOL_NB, tmp, tmp2);
    VMCX_LAUNCHED();
  }
  // backward: x = L^-T tmp2
  for (int64_t kb = nblk - 1; kb >= 0; --kb) {
    const int64_t k0 = kb * CHOL_NB;
    const int nbk = (int)((n - k0 < CHOL_NB) ? n - k0 : CHOL_NB);
    const unsigned grid = (unsigned)((k0 + 3) / 4 > 0 ? (k0 + 3) / 4 : 1);
    hipLaunchKernelGGL(trsv_bwd_kernel, dim3(grid), dim3(256), 0, st, L, lda, k0, nbk,
                       DinvT + kb * CHOL_NB * CHOL_NB, tmp2, x);
    VMCX_LAUNCHED();
  }
  return 0;
}

}  // namespace vmcx
