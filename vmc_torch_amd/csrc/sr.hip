// Stochastic reconfiguration in sample space (minSR) on gfx950.
//
//   Ot  = (O - mean_s O) / sqrt(Ns)          (Ns x Np, row-major = X, Np x Ns
//                                            column-major; X[:, s] = Ot[s, :])
//   eps = (E - E_mean) / sqrt(Ns)
//   T   = Ot Ot^T + shift I                  Gram in FP64 (SYRK), or FP32 /
//                                            BF16 operands with FP32 accumulate
//   L L^T = T (FP64 Cholesky), alpha = T^-1 eps, refined against the FP64
//   operator when the Gram was formed in reduced precision
//   dp  = Ot^T alpha
//
// Column means and the energy moments use Neumaier-compensated sums.  The
// FP64 Gram is the matrix-core SYRK kernel (syrk.hip), the Cholesky and its
// solves are chol.hip, the GEMV passes over Ot and the fused centring /
// conversion / shift kernels are below; rocBLAS serves the reduced-precision
// Gram (gemm_ex) and, in the factored form at the end of this file, the
// per-block update GEMMs (dgemm) and the symmetric products with T (dsymv).
#include <hip/hip_bf16.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <map>
#include <unordered_map>
#include <vector>

#include "kernels.h"

namespace vmcx {

// ------------------------------------------------------------ reductions
constexpr int kColsumChunks = 64;   // row chunks of the column sums

struct Neu {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

// partial[r][p] = sum over rows in chunk r of w[s] O[s][p] (w = NULL: 1);
// with w2 != NULL also partial2[r][p] = sum of w2[s] O[s][p] from the same
// read of O.  skip != NULL and *skip != 0: the launch does nothing (the
// device-side convergence flag of the conjugate-gradient loop).
template <typename T>
__global__ void __launch_bounds__(256)
    colsum_partial_kernel(int64_t Ns, int64_t Np, const T *__restrict__ O, int64_t ldo,
                          int64_t rows_per_chunk, const double *__restrict__ w,
                          double *partial, const double *__restrict__ w2, double *partial2,
                          const double *skip) {
  if (skip && skip[0] != 0.0) return;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
  const int64_t r1 = (r0 + rows_per_chunk < Ns) ? r0 + rows_per_chunk : Ns;
  Neu acc, acc2;
  if (w2) {
    for (int64_t s = r0; s < r1; ++s) {
      const double v = (double)O[s * ldo + p];
      acc.add(w ? v * w[s] : v);
      acc2.add(v * w2[s]);
    }
    partial2[(int64_t)blockIdx.y * Np + p] = acc2.value();
  } else {
    for (int64_t s = r0; s < r1; ++s) {
      const double v = (double)O[s * ldo + p];
      acc.add(w ? v * w[s] : v);
    }
  }
  partial[(int64_t)blockIdx.y * Np + p] = acc.value();
}

__global__ void __launch_bounds__(256)
    colsum_final_kernel(int64_t Np, int R, const double *__restrict__ partial,
                        double scale, double *out, const double *skip) {
  if (skip && skip[0] != 0.0) return;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  Neu acc;
  for (int r = 0; r < R; ++r) acc.add(partial[(int64_t)r * Np + p]);
  out[p] = acc.value() * scale;
}

// v[s] = sum_p X[s][p] u[p]   (one block per row, FP64 accumulation)
template <typename T>
__global__ void __launch_bounds__(256)
    gemv_rows_kernel(int64_t Ns, int64_t Np, const T *__restrict__ X, int64_t ldx,
                     const double *__restrict__ u, double *v, const double *skip) {
  if (skip && skip[0] != 0.0) return;
  __shared__ double sh[4];
  const int64_t s = blockIdx.x;
  const T *row = X + s * ldx;
  double acc = 0.0;
  for (int64_t p = threadIdx.x; p < Np; p += blockDim.x) acc += (double)row[p] * u[p];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) v[s] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// X[s][p] = (O[s][p] - mean[p]) * scale (X may alias O); optional low copy
// Xlo[s][p] (row stride Np).  Thread = one column p, its mean in a register;
// the workgroups of one column block stride through the rows (blockIdx.y),
// so every load and store of a wave is 64 consecutive elements.
template <typename T, typename TL>
__global__ void __launch_bounds__(256)
    center_kernel(int64_t Ns, int64_t Np, const T *O, int64_t ldo,
                  const double *__restrict__ mean, double scale, T *X, int64_t ldx,
                  TL *Xlo) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  const double m = mean[p];
#pragma unroll 4
  for (int64_t s = blockIdx.y; s < Ns; s += gridDim.y) {
    const double v = ((double)O[s * ldo + p] - m) * scale;
    X[s * ldx + p] = (T)v;
    if (Xlo) Xlo[s * Np + p] = (TL)v;
  }
}

// Split form of the centred matrix for the three-product BF16 Gram:
// X[s][p] = v (v = (O - mean) * scale, may alias O), Xhi = bf16(v),
// Xlo = bf16(v - Xhi), so Xhi + Xlo carries v to ~2^-16.
template <typename T>
__global__ void __launch_bounds__(256)
    center_split_kernel(int64_t Ns, int64_t Np, const T *O, int64_t ldo,
                        const double *__restrict__ mean, double scale, T *X, int64_t ldx,
                        __hip_bfloat16 *Xhi, __hip_bfloat16 *Xlo) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  const double m = mean[p];
#pragma unroll 4
  for (int64_t s = blockIdx.y; s < Ns; s += gridDim.y) {
    const double v = ((double)O[s * ldo + p] - m) * scale;
    X[s * ldx + p] = (T)v;
    const __hip_bfloat16 hi = __float2bfloat16((float)v);
    Xhi[s * Np + p] = hi;
    Xlo[s * Np + p] = __float2bfloat16((float)(v - (double)__bfloat162float(hi)));
  }
}

// T (column-major, lower used) <- G + H + H^T (+ shift on the diagonal):
// the three-product Gram hi hi^T + hi lo^T + lo hi^T from G = hi hi^T and
// H = hi lo^T (FP32, column-major Ns x Ns)
__global__ void __launch_bounds__(256)
    shift3_kernel(int64_t Ns, const float *__restrict__ G, const float *__restrict__ H,
                  double *T, double shift) {
  const int64_t n = Ns * Ns;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx / Ns, i = idx - j * Ns;
    double v = (double)G[idx] + (double)H[idx] + (double)H[i * Ns + j];
    if (i == j) v += shift;
    T[idx] = v;
  }
}

// One block: mean and variance of E (Neumaier), eps = (E - Em) * scale.
// stats: [0] mean, [1] var (<E^2> - <E>^2)
__global__ void __launch_bounds__(1024)
    energy_kernel(int64_t Ns, const double *__restrict__ E, double given_mean,
                  int use_given, double scale, double *eps, double *stats) {
  __shared__ double sh[4][32];
  Neu a, b;
  for (int64_t s = threadIdx.x; s < Ns; s += blockDim.x) {
    a.add(E[s]);
    b.add(E[s] * E[s]);
  }
  double v0 = a.s, v1 = a.c, v2 = b.s, v3 = b.c;
  // wave-level tree (values stay compensated pairs combined by Neumaier)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) {
    Neu x, y;
    x.s = v0; x.c = v1; x.add(__shfl_down(v0, o, 64)); x.c += __shfl_down(v1, o, 64);
    y.s = v2; y.c = v3; y.add(__shfl_down(v2, o, 64)); y.c += __shfl_down(v3, o, 64);
    v0 = x.s; v1 = x.c; v2 = y.s; v3 = y.c;
  }
  if (lane == 0) {
    sh[0][wv] = v0; sh[1][wv] = v1; sh[2][wv] = v2; sh[3][wv] = v3;
  }
  __syncthreads();
  __shared__ double mean_sh;
  if (threadIdx.x == 0) {
    Neu x, y;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
      x.add(sh[0][k]); x.c += sh[1][k];
      y.add(sh[2][k]); y.c += sh[3][k];
    }
    const double m = x.value() / (double)Ns;
    const double m2 = y.value() / (double)Ns;
    stats[0] = m;
    stats[1] = m2 - m * m;
    mean_sh = use_given ? given_mean : m;
  }
  __syncthreads();
  const double em = mean_sh;
  for (int64_t s = threadIdx.x; s < Ns; s += blockDim.x) eps[s] = (E[s] - em) * scale;
}

// stats <- [mean E, var E, ||res|| / ||eps||, ||g||, PCG iterations]
__global__ void finish_stats(const double *sc, double *stats) {
  stats[0] = sc[8];
  stats[1] = sc[9];
  stats[2] = sc[11] > 0.0 ? sc[10] / sc[11] : sc[10];
  stats[3] = sc[12];
  stats[4] = sc[6];
}

// T (column-major Ns x Ns, lower used) <- Tin (+ shift on the diagonal)
template <typename TI>
__global__ void __launch_bounds__(256)
    shift_kernel(int64_t Ns, const TI *__restrict__ Tin, double *T, double shift) {
  const int64_t n = Ns * Ns;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx / Ns, i = idx - j * Ns;
    double v = (double)Tin[idx];
    if (i == j) v += shift;
    T[idx] = v;
  }
}

// T += sum over the partial triangles (lower, column-major) + shift on the
// diagonal
__global__ void __launch_bounds__(256)
    add_parts_kernel(int64_t Ns, int parts, const double *__restrict__ Tp, double *T,
                     double shift) {
  const int64_t n = Ns * Ns;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx / Ns, i = idx - j * Ns;
    if (i < j) continue;
    double v = T[idx];
    for (int p = 0; p < parts; ++p) v += Tp[(int64_t)p * n + idx];
    if (i == j) v += shift;
    T[idx] = v;
  }
}

__global__ void diag_shift_kernel(int64_t Ns, double *T, double shift) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < Ns) T[i * Ns + i] += shift;
}

// r = eps - v - shift * alpha
__global__ void residual_kernel(int64_t Ns, const double *eps, const double *v,
                                const double *alpha, double shift, double *r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < Ns) r[i] = eps[i] - v[i] - shift * alpha[i];
}

__global__ void axpy_kernel(int64_t n, double a, const double *x, double *y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

__global__ void copy_kernel(int64_t n, const double *x, double *y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

// sum of squares (one block) -> out[0] (Neumaier)
__global__ void __launch_bounds__(1024)
    norm2_kernel(int64_t n, const double *x, double *out) {
  __shared__ double sh[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += x[i] * x[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    Neu a;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) a.add(sh[k]);
    out[0] = sqrt(a.value());
  }
}

// --------------------------------------------------------------- handles
// One rocBLAS handle per (device, caller stream): a handle is bound to the
// stream it was created for, so callers on different streams never share
// one (and never see another caller's stream or pointer mode).
rocblas_handle blas_handle_for(hipStream_t st) {
  static std::mutex mu;
  static std::map<std::pair<int, hipStream_t>, rocblas_handle> handles;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  const auto key = std::make_pair(dev, st);
  auto it = handles.find(key);
  if (it != handles.end()) return it->second;
  rocblas_handle h;
  if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
  rocblas_set_stream(h, st);
  rocblas_set_pointer_mode(h, rocblas_pointer_mode_host);
  handles[key] = h;
  return h;
}

#define VMCX_BLAS(call)                                                       \
  do {                                                                       \
    rocblas_status s_ = (call);                                              \
    if (s_ != rocblas_status_success) {                                      \
      ::vmcx::set_error(std::string(__func__) + ": " #call " -> " +          \
                        rocblas_status_to_string(s_));                       \
      return 3;                                                              \
    }                                                                        \
  } while (0)

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct SrLayout {
  size_t X, Xlo, Xlo2, T, Tlo, Tlo2, Tpart, partial, mean, eps, alpha, r, z, pv, q, u, stw, info, dinv, t1,
      t2, total;
};

// FP64 Gram on the matrix-core SYRK (syrk.hip, one workgroup per 128 x 128
// lower tile, two per CU): when the triangle has too few tiles to fill the
// chip a few times over, the reduction is split; split 0 writes T itself and
// splits 1.. write partial triangles that one pass adds into T.
// Splits until ~3000 tiles (about six waves of 512 workgroups), at most 32,
// chunks of at least 1024 parameters.
static int gram_splits(int64_t Ns, int64_t Np) {
  const int64_t t = (Ns + 127) / 128;
  const int64_t tiles = t * (t + 1) / 2;
  int64_t s = (3000 + tiles - 1) / tiles;
  if (s > 32) s = 32;
  if (s > Np / 1024) s = Np / 1024;
  return s < 2 ? 1 : (int)s;
}

static SrLayout sr_layout(int64_t Ns, int64_t Np, int o_dtype, int gram, int inplace) {
  SrLayout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  const int R = kColsumChunks;
  const size_t osz = o_dtype == 0 ? 8 : 4;
  L.X = take(inplace ? 0 : (size_t)Ns * Np * osz);
  const bool need_lo = (gram >= 2) || (gram == 1 && o_dtype == 0);
  L.Xlo = take(need_lo ? (size_t)Ns * Np * (gram >= 2 ? 2 : 4) : 0);
  L.Xlo2 = take(gram == 3 ? (size_t)Ns * Np * 2 : 0);
  L.T = take((size_t)Ns * Ns * 8);
  L.Tlo = take(gram ? (size_t)Ns * Ns * 4 : 0);
  L.Tlo2 = take(gram == 3 ? (size_t)Ns * Ns * 4 : 0);
  const int ns = gram == 0 ? gram_splits(Ns, Np) : 1;
  L.Tpart = take(ns > 1 ? (size_t)(ns - 1) * Ns * Ns * 8 : 0);
  L.partial = take((size_t)2 * R * Np * 8);
  L.mean = take((size_t)Np * 8);
  L.eps = take((size_t)Ns * 8);
  L.alpha = take((size_t)Ns * 8);
  L.r = take((size_t)Ns * 8);
  L.z = take((size_t)Ns * 8);
  L.pv = take((size_t)Ns * 8);
  L.q = take((size_t)Ns * 8);
  L.u = take((size_t)Np * 8);
  L.stw = take(16 * 8);
  L.info = take(64);
  L.dinv = take((size_t)chol_dinv_bytes(Ns));
  L.t1 = take((size_t)Ns * 8);
  L.t2 = take((size_t)Ns * 8);
  L.total = off;
  return L;
}

int64_t sr_workspace_bytes(int64_t Ns, int64_t Np, int o_dtype, int gram, int flags) {
  if (Ns <= 0 || Np <= 0 || gram < 0 || gram > 3 || o_dtype < 0 || o_dtype > 1) return -1;
  return (int64_t)sr_layout(Ns, Np, o_dtype, gram, flags & 1).total;
}

// ---- small device-side scalar kernels for PCG (no host round trips)
// dot(a, b) -> out[0] (one block, Neumaier)
__global__ void __launch_bounds__(1024)
    dot_kernel(int64_t n, const double *a, const double *b, double *out) {
  __shared__ double sh[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += a[i] * b[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    Neu acc;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) acc.add(sh[k]);
    out[0] = acc.value();
  }
}

// scalars: [0] rz, [1] pq, [2] rz_new, [3] |r|^2, [4] |eps|^2, [5] done, [6] iters
// x += a p ; r -= a q  with a = rz / pq   (skipped once done)
__global__ void pcg_update_xr(int64_t n, double *x, double *r, const double *p,
                              const double *q, const double *sc) {
  if (sc[5] != 0.0) return;
  const double a = sc[0] / sc[1];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    x[i] += a * p[i];
    r[i] -= a * q[i];
  }
}

// check convergence after the x/r update: done if |r| <= tol |eps|
__global__ void pcg_check(double tol, double *sc) {
  if (sc[5] != 0.0) return;
  sc[6] += 1.0;
  if (sc[3] <= tol * tol * sc[4]) sc[5] = 1.0;
}

// dp += a Ot^T p  (u = Ot^T p), same a as pcg_update_xr; skipped once done
__global__ void pcg_update_d(int64_t n, double *d, const double *u, const double *sc) {
  if (sc[5] != 0.0) return;
  const double a = sc[0] / sc[1];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] += a * u[i];
}

// the direct solve may already meet the tolerance: no iteration then
__global__ void pcg_check0(double tol, double *sc) {
  if (sc[3] <= tol * tol * sc[4]) sc[5] = 1.0;
}

// p = z + (rz_new / rz) p ; rz = rz_new
__global__ void pcg_update_p(int64_t n, double *p, const double *z, double *sc) {
  if (sc[5] != 0.0) return;
  const double b = sc[2] / sc[0];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + b * p[i];
}

__global__ void pcg_rotate(double *sc) {
  if (sc[5] != 0.0) return;
  sc[0] = sc[2];
}

// r = eps - v - shift * x
__global__ void residual2_kernel(int64_t n, const double *eps, const double *v,
                                 const double *x, double shift, double *r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = eps[i] - v[i] - shift * x[i];
}

// q = v + shift p
__global__ void addshift_kernel(int64_t n, const double *v, const double *p, double shift,
                                double *q) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) q[i] = v[i] + shift * p[i];
}

__global__ void set_scalars(double *sc) {
  for (int i = 0; i < 16; ++i) sc[i] = 0.0;
}

__global__ void copy_scalar(const double *src, double *dst) { dst[0] = src[0]; }

template <typename T>
static int xt_gemv_rows(int64_t Ns, int64_t Np, const T *X, int64_t ldx, const double *u,
                        double *v, hipStream_t st, const double *skip = nullptr) {
  hipLaunchKernelGGL(gemv_rows_kernel<T>, dim3((unsigned)Ns), dim3(256), 0, st, Ns, Np, X, ldx,
                     u, v, skip);
  VMCX_LAUNCHED();
  return 0;
}

// out[p] = scale * sum_s X[s][p] w[s]; with w2: out2[p] = scale * sum_s X[s][p] w2[s]
// (partial holds 2 * kColsumChunks * Np doubles)
template <typename T>
static int xt_gemv_cols(int64_t Ns, int64_t Np, const T *X, int64_t ldx, const double *w,
                        double scale, double *partial, double *out, hipStream_t st,
                        const double *skip = nullptr, const double *w2 = nullptr,
                        double *out2 = nullptr) {
  const int R = kColsumChunks;
  const int64_t rows = (Ns + R - 1) / R;
  double *partial2 = partial + (int64_t)R * Np;
  const unsigned gx = (unsigned)((Np + 255) / 256);
  hipLaunchKernelGGL(colsum_partial_kernel<T>, dim3(gx, R), dim3(256), 0, st, Ns, Np, X, ldx,
                     rows, w, partial, w2, partial2, skip);
  VMCX_LAUNCHED();
  hipLaunchKernelGGL(colsum_final_kernel, dim3(gx), dim3(256), 0, st, Np, R, partial, scale, out,
                     skip);
  VMCX_LAUNCHED();
  if (w2) {
    hipLaunchKernelGGL(colsum_final_kernel, dim3(gx), dim3(256), 0, st, Np, R, partial2, scale,
                       out2, skip);
    VMCX_LAUNCHED();
  }
  return 0;
}

// Optional phase timing: HIP events on the launch stream; read back with one
// host synchronisation at the end of the call (only when phase_ms != NULL).
struct PhaseTimer {
  hipEvent_t ev[6];
  int n = 0;
  bool on;
  hipStream_t st;
  PhaseTimer(bool enable, hipStream_t s) : on(enable), st(s) {
    if (on)
      for (auto &e : ev) (void)hipEventCreate(&e);
  }
  void mark() {
    if (on && n < 6) (void)hipEventRecord(ev[n++], st);
  }
  void read(float *out) {
    if (!on) return;
    (void)hipEventSynchronize(ev[n - 1]);
    for (int i = 0; i + 1 < n; ++i) (void)hipEventElapsedTime(&out[i], ev[i], ev[i + 1]);
  }
  ~PhaseTimer() {
    if (on)
      for (auto &e : ev) (void)hipEventDestroy(e);
  }
};

template <typename T>
static int sr_minsr_t(int64_t Ns, int64_t Np, T *O, int64_t ldo, int flags, const double *E,
                      double energy_mean, double shift, int gram, int max_iter, double tol,
                      char *w, const SrLayout &L, double *dp, double *stats, int32_t *info,
                      float *phase_ms, hipStream_t st) {
  const bool inplace = flags & 1;
  T *X = inplace ? O : (T *)(w + L.X);
  const int64_t ldx = inplace ? ldo : Np;
  double *T64 = (double *)(w + L.T);
  double *partial = (double *)(w + L.partial), *mean = (double *)(w + L.mean);
  double *eps = (double *)(w + L.eps), *alpha = (double *)(w + L.alpha);
  double *r = (double *)(w + L.r), *z = (double *)(w + L.z), *pv = (double *)(w + L.pv);
  double *q = (double *)(w + L.q), *u = (double *)(w + L.u), *sc = (double *)(w + L.stw);
  double *dinv = (double *)(w + L.dinv), *t1 = (double *)(w + L.t1), *t2 = (double *)(w + L.t2);
  int32_t *iinfo = (int32_t *)(w + L.info);
  const double scale = 1.0 / std::sqrt((double)Ns);
  const int nsi = (int)Ns, npi = (int)Np;
  rocblas_handle h = blas_handle_for(st);
  VMCX_CHECK_ARG(h != nullptr, "rocblas_create_handle failed");
  PhaseTimer timer(phase_ms != nullptr, st);
  timer.mark();
  int rc;

  // column means (two-level compensated sums) and centring
  rc = xt_gemv_cols<T>(Ns, Np, O, ldo, nullptr, 1.0 / (double)Ns, partial, mean, st);
  if (rc) return rc;
  const unsigned cgx = (unsigned)((Np + 255) / 256);
  const dim3 cgrid(cgx, (unsigned)std::max<int64_t>(1, std::min<int64_t>(Ns, 16384 / cgx)));
  if (gram == 3)
    hipLaunchKernelGGL(center_split_kernel<T>, cgrid, dim3(256), 0, st, Ns, Np, O, ldo, mean,
                       scale, X, ldx, (__hip_bfloat16 *)(w + L.Xlo),
                       (__hip_bfloat16 *)(w + L.Xlo2));
  else if (gram == 2)
    hipLaunchKernelGGL((center_kernel<T, __hip_bfloat16>), cgrid, dim3(256), 0, st, Ns,
                       Np, O, ldo, mean, scale, X, ldx, (__hip_bfloat16 *)(w + L.Xlo));
  else if (gram == 1 && sizeof(T) == 8)
    hipLaunchKernelGGL((center_kernel<T, float>), cgrid, dim3(256), 0, st, Ns, Np, O, ldo,
                       mean, scale, X, ldx, (float *)(w + L.Xlo));
  else
    hipLaunchKernelGGL((center_kernel<T, float>), cgrid, dim3(256), 0, st, Ns, Np, O, ldo,
                       mean, scale, X, ldx, (float *)nullptr);
  VMCX_LAUNCHED();
  const int use_given = std::isfinite(energy_mean) ? 1 : 0;
  hipLaunchKernelGGL(set_scalars, dim3(1), dim3(1), 0, st, sc);
  VMCX_LAUNCHED();
  hipLaunchKernelGGL(energy_kernel, dim3(1), dim3(1024), 0, st, Ns, E, energy_mean, use_given,
                     scale, eps, sc + 8);
  VMCX_LAUNCHED();
  timer.mark();

  // Gram
  const double one = 1.0, zero = 0.0;
  if (gram == 0) {
    const int ns = gram_splits(Ns, Np);
    double *Tp = (double *)(w + L.Tpart);
    rc = syrk_f64(Ns, Np, (const double *)X, ldx, 1, 1.0, 0.0, T64, Ns, ns, Ns * Ns, st, Tp);
    if (rc) return rc;
    if (ns == 1)
      hipLaunchKernelGGL(diag_shift_kernel, dim3((unsigned)((Ns + 255) / 256)), dim3(256), 0,
                         st, Ns, T64, shift);
    else
      hipLaunchKernelGGL(add_parts_kernel, dim3(4096), dim3(256), 0, st, Ns, ns - 1, Tp, T64,
                         shift);
    VMCX_LAUNCHED();
  } else {
    const float onef = 1.0f, zerof = 0.0f;
    const bool direct32 = (gram == 1 && sizeof(T) == 4);
    const rocblas_datatype ab = (gram == 1) ? rocblas_datatype_f32_r : rocblas_datatype_bf16_r;
    const void *Xg = direct32 ? (const void *)X : (const void *)(w + L.Xlo);
    const int ldg = direct32 ? (int)ldx : npi;
    float *Tlo = (float *)(w + L.Tlo);
    VMCX_BLAS(rocblas_gemm_ex(h, rocblas_operation_transpose, rocblas_operation_none, nsi, nsi,
                              npi, &onef, Xg, ab, ldg, Xg, ab, ldg, &zerof, Tlo,
                              rocblas_datatype_f32_r, nsi, Tlo, rocblas_datatype_f32_r, nsi,
                              rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0));
    if (gram == 3) {
      // hi lo^T; the Gram is hi hi^T + hi lo^T + (hi lo^T)^T
      float *Tlo2 = (float *)(w + L.Tlo2);
      VMCX_BLAS(rocblas_gemm_ex(h, rocblas_operation_transpose, rocblas_operation_none, nsi,
                                nsi, npi, &onef, Xg, ab, ldg, (const void *)(w + L.Xlo2), ab,
                                ldg, &zerof, Tlo2, rocblas_datatype_f32_r, nsi, Tlo2,
                                rocblas_datatype_f32_r, nsi, rocblas_datatype_f32_r,
                                rocblas_gemm_algo_standard, 0, 0));
      hipLaunchKernelGGL(shift3_kernel, dim3(4096), dim3(256), 0, st, Ns, Tlo, Tlo2, T64,
                         shift);
    } else {
      hipLaunchKernelGGL(shift_kernel<float>, dim3(4096), dim3(256), 0, st, Ns, Tlo, T64,
                         shift);
    }
    VMCX_LAUNCHED();
  }
  timer.mark();
  // FP64 Cholesky of the (shifted) Gram
  // (the forward half of the first solve, L^-1 eps, runs along the factorisation)
  VMCX_HIP(hipMemcpyAsync(t1, eps, Ns * sizeof(double), hipMemcpyDeviceToDevice, st));
  rc = chol_factor((void *)h, Ns, T64, Ns, dinv, iinfo, st, phase_ms ? phase_ms + 5 : nullptr,
                   t1, t2);
  if (rc) return rc;
  if (gram != 0 && max_iter > 0) {
    // The reduced-precision Gram only preconditions CG on the exact operator.
    // Its rounding (~2^-8 |T| in BF16) can leave it indefinite at the given
    // shift; then the preconditioner's shift is raised tenfold until the
    // factorisation succeeds (up to 10^6 x), the operator keeps `shift`.
    double pc_shift = shift;
    for (int attempt = 0; attempt < 6; ++attempt) {
      int32_t hinfo = 0;
      VMCX_HIP(hipMemcpyAsync(&hinfo, iinfo, sizeof(int32_t), hipMemcpyDeviceToHost, st));
      VMCX_HIP(hipStreamSynchronize(st));
      if (hinfo == 0) break;
      pc_shift = (pc_shift > 0.0 ? pc_shift : 1e-8) * 10.0;
      if (gram == 3)
        hipLaunchKernelGGL(shift3_kernel, dim3(4096), dim3(256), 0, st, Ns,
                           (const float *)(w + L.Tlo), (const float *)(w + L.Tlo2), T64,
                           pc_shift);
      else
        hipLaunchKernelGGL(shift_kernel<float>, dim3(4096), dim3(256), 0, st, Ns,
                           (const float *)(w + L.Tlo), T64, pc_shift);
      VMCX_LAUNCHED();
      VMCX_HIP(hipMemcpyAsync(t1, eps, Ns * sizeof(double), hipMemcpyDeviceToDevice, st));
      rc = chol_factor((void *)h, Ns, T64, Ns, dinv, iinfo, st, nullptr, t1, t2);
      if (rc) return rc;
    }
  }
  timer.mark();
  rc = chol_solve_bwd(Ns, chol_panels(Ns, dinv), Ns, dinv, t2, alpha, st);
  if (rc) return rc;
  // dp = Ot^T alpha and u = Ot^T eps (= the energy gradient, for ||g||) from
  // one read of Ot
  rc = xt_gemv_cols<T>(Ns, Np, X, ldx, alpha, 1.0, partial, dp, st, nullptr, eps, u);
  if (rc) return rc;
  hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, st, Np, u, sc + 12);
  VMCX_LAUNCHED();
  // preconditioned CG on (Ot Ot^T + shift) alpha = eps, preconditioner = the
  // factor.  dp = Ot^T alpha is carried along (dp += a Ot^T p), so no pass over
  // Ot is spent on it after the loop.  Once converged (sc[5] set on the
  // device) every remaining launch returns at once.
  const unsigned gv = (unsigned)((Ns + 255) / 256);
  const unsigned gp = (unsigned)((Np + 255) / 256);
  const double *done = sc + 5;
  if (max_iter > 0) {
    rc = xt_gemv_rows<T>(Ns, Np, X, ldx, dp, q, st);
    if (rc) return rc;
    hipLaunchKernelGGL(residual2_kernel, dim3(gv), dim3(256), 0, st, Ns, eps, q, alpha, shift, r);
    VMCX_LAUNCHED();
    rc = chol_solve(Ns, chol_panels(Ns, dinv), Ns, dinv, r, z, t1, t2, st);
    if (rc) return rc;
    VMCX_HIP(hipMemcpyAsync(pv, z, Ns * 8, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, st, Ns, r, z, sc + 0);
    hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, st, Ns, eps, eps, sc + 4);
    hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, st, Ns, r, r, sc + 3);
    hipLaunchKernelGGL(pcg_check0, dim3(1), dim3(1), 0, st, tol, sc);
    VMCX_LAUNCHED();
    for (int it = 0; it < max_iter; ++it) {
      rc = xt_gemv_cols<T>(Ns, Np, X, ldx, pv, 1.0, partial, u, st, done);
      if (rc) return rc;
      rc = xt_gemv_rows<T>(Ns, Np, X, ldx, u, t1, st, done);
      if (rc) return rc;
      hipLaunchKernelGGL(addshift_kernel, dim3(gv), dim3(256), 0, st, Ns, t1, pv, shift, q);
      hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, st, Ns, pv, q, sc + 1);
      hipLaunchKernelGGL(pcg_update_xr, dim3(gv), dim3(256), 0, st, Ns, alpha, r, pv, q, sc);
      hipLaunchKernelGGL(pcg_update_d, dim3(gp), dim3(256), 0, st, Np, dp, u, sc);
      hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, st, Ns, r, r, sc + 3);
      hipLaunchKernelGGL(pcg_check, dim3(1), dim3(1), 0, st, tol, sc);
      VMCX_LAUNCHED();
      if (it + 1 == max_iter) break;
      rc = chol_solve(Ns, chol_panels(Ns, dinv), Ns, dinv, r, z, t1, t2, st, done);
      if (rc) return rc;
      hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, st, Ns, r, z, sc + 2);
      hipLaunchKernelGGL(pcg_update_p, dim3(gv), dim3(256), 0, st, Ns, pv, z, sc);
      hipLaunchKernelGGL(pcg_rotate, dim3(1), dim3(1), 0, st, sc);
      VMCX_LAUNCHED();
    }
  }
  timer.mark();
  if (stats) {
    // explicit residual of the final alpha: eps - Ot dp - shift alpha
    rc = xt_gemv_rows<T>(Ns, Np, X, ldx, dp, q, st);
    if (rc) return rc;
    hipLaunchKernelGGL(residual2_kernel, dim3(gv), dim3(256), 0, st, Ns, eps, q, alpha, shift, r);
    hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, st, Ns, r, sc + 10);
    hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, st, Ns, eps, sc + 11);
    VMCX_LAUNCHED();
    // stats: mean, var, ||res||/||eps||, ||g||, iterations
    hipLaunchKernelGGL(finish_stats, dim3(1), dim3(1), 0, st, sc, stats);
    VMCX_LAUNCHED();
  }
  if (info) VMCX_HIP(hipMemcpyAsync(info, iinfo, sizeof(int32_t), hipMemcpyDeviceToDevice, st));
  timer.mark();
  timer.read(phase_ms);
  return 0;
}

int sr_minsr(int64_t Ns, int64_t Np, void *O, int o_dtype, int64_t ldo, int flags,
             const double *E, double energy_mean, double shift, int gram, int max_iter,
             double tol, void *ws, int64_t ws_bytes, double *dp, double *stats, int32_t *info,
             float *phase_ms, hipStream_t st) {
  VMCX_CHECK_ARG(Ns >= 1 && Np >= 1, "Ns and Np must be positive");
  VMCX_CHECK_ARG(Ns <= 0x7fffffff && Np <= 0x7fffffff, "Ns / Np exceed 32-bit BLAS");
  VMCX_CHECK_ARG(ldo >= Np, "ldo < Np");
  VMCX_CHECK_ARG(o_dtype == 0 || o_dtype == 1, "o_dtype must be 0 (f64) or 1 (f32)");
  VMCX_CHECK_ARG(gram >= 0 && gram <= 3,
                 "gram must be 0 (f64), 1 (f32), 2 (bf16) or 3 (three-product bf16)");
  VMCX_CHECK_ARG(!(gram == 0 && o_dtype == 1), "an FP64 Gram needs FP64 O");
  VMCX_CHECK_ARG(max_iter >= 0 && max_iter <= 1000, "max_iter must be in [0, 1000]");
  VMCX_CHECK_ARG(O && E && ws && dp, "null pointer");
  const SrLayout L = sr_layout(Ns, Np, o_dtype, gram, flags & 1);
  VMCX_CHECK_ARG(ws_bytes >= (int64_t)L.total, "workspace too small");
  if (o_dtype == 0)
    return sr_minsr_t<double>(Ns, Np, (double *)O, ldo, flags, E, energy_mean, shift, gram,
                              max_iter, tol, (char *)ws, L, dp, stats, info, phase_ms, st);
  return sr_minsr_t<float>(Ns, Np, (float *)O, ldo, flags, E, energy_mean, shift, gram,
                           max_iter, tol, (char *)ws, L, dp, stats, info, phase_ms, st);
}

// ===================================================== minSR on a factored O
// O = [blocks], each a factor F_a or an outer product scale F_a (x) F_b (+ a
// bias block scale F_a).  Row s of an outer-product block is the flattened
// F_a[s] F_b[s]^T, so
//   (O O^T)[s][t] = sum_b scale_b^2 (F_a[s] . F_a[t]) (F_b[s] . F_b[t] [+ 1])
// (Kronecker identity; the + 1 is the bias row).  The uncentred Gram K comes
// from the factor Grams G_f = F_f F_f^T (SYRK, inner dimension k_f), the
// centred one is P K P (P = I - 1 1^T / Ns): Ot Ot^T = P K P / Ns.  The
// update dp = Ot^T alpha = O^T (P alpha) / sqrt(Ns) is, per outer-product
// block, F_a^T diag(a~) F_b: one FP64 GEMM with inner dimension Ns.
constexpr int kFactMax = 8;

struct GramTerms {
  const double *ga[kFactMax];
  const double *gb[kFactMax];
  double s2[kFactMax];
  double bias[kFactMax];
  int n;
};

// K (column-major, lower triangle) = sum_b s2_b G_a o (G_b + bias_b); G_b
// NULL: s2_b G_a
__global__ void __launch_bounds__(256)
    gram_combine_kernel(int64_t Ns, GramTerms t, double *K) {
  const int64_t n = Ns * Ns;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx / Ns, i = idx - j * Ns;
    if (i < j) continue;
    double v = 0.0;
    for (int b = 0; b < t.n; ++b) {
      const double g = t.gb[b] ? t.gb[b][idx] + t.bias[b] : 1.0;
      v += t.s2[b] * t.ga[b][idx] * g;
    }
    K[idx] = v;
  }
}

// T = (K - r 1^T / Ns - 1 r^T / Ns + sum(r) / Ns^2) / Ns + shift I on the
// lower triangle (in place; r = K 1, *rsum = sum of r)
__global__ void __launch_bounds__(256)
    double_center_kernel(int64_t Ns, const double *__restrict__ r, const double *rsum,
                         double shift, double *T) {
  const double inv = 1.0 / (double)Ns;
  const double mu = rsum[0] * inv * inv;
  const int64_t n = Ns * Ns;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx / Ns, i = idx - j * Ns;
    if (i < j) continue;
    double v = (T[idx] - r[i] * inv - r[j] * inv + mu) * inv;
    if (i == j) v += shift;
    T[idx] = v;
  }
}

__global__ void fill_kernel(int64_t n, double v, double *x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// out = (a - sum(a) / n) * scale, sum(a) in *asum
__global__ void center_vec_kernel(int64_t n, const double *a, const double *asum, double scale,
                                  double *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (a[i] - asum[0] / (double)n) * scale;
}

// Y[s][j] = w[s] F[s][j]  (Y row-major [Ns][kb])
__global__ void __launch_bounds__(256)
    scale_rows_kernel(int64_t Ns, int kb, const double *__restrict__ F, int64_t ld,
                      const double *__restrict__ w, double *Y) {
  const int64_t n = Ns * kb;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = idx / kb, j = idx - s * kb;
    Y[idx] = w[s] * F[s * ld + j];
  }
}

// sc[12] = ||g|| = sqrt(eps^T (T - shift I) eps) from sc[14] = eps^T T eps and
// sc[4] = eps^T eps
__global__ void gnorm_kernel(double shift, double *sc) {
  sc[12] = sqrt(fmax(0.0, sc[14] - shift * sc[4]));
}

struct FactLayout {
  size_t G, Gp, T, Tc, dinv, vec, sc, info, Y, part, total;
  int splits[kFactMax];
};

static int check_factored(int64_t Ns, int n_factors, const int32_t *k, int n_blocks,
                          const vmcx_sr_block *blocks, int64_t Np) {
  VMCX_CHECK_ARG(Ns >= 1 && Ns <= 0x7fffffff, "Ns must be in [1, 2^31)");
  VMCX_CHECK_ARG(n_factors >= 1 && n_factors <= kFactMax, "n_factors must be in [1, 8]");
  VMCX_CHECK_ARG(n_blocks >= 1 && n_blocks <= kFactMax, "n_blocks must be in [1, 8]");
  VMCX_CHECK_ARG(k && blocks, "null pointer");
  for (int f = 0; f < n_factors; ++f) VMCX_CHECK_ARG(k[f] >= 1, "factor width < 1");
  if (Np < 0) return 0;   // workspace query: no coverage check
  // the blocks must tile [0, Np)
  std::vector<std::pair<int64_t, int64_t>> iv;
  for (int b = 0; b < n_blocks; ++b) {
    const vmcx_sr_block &B = blocks[b];
    VMCX_CHECK_ARG(B.a_factor >= 0 && B.a_factor < n_factors, "a_factor out of range");
    VMCX_CHECK_ARG(B.b_factor >= -1 && B.b_factor < n_factors, "b_factor out of range");
    VMCX_CHECK_ARG(B.b_factor >= 0 || B.bias_offset < 0, "a plain block has no bias");
    const int64_t ka = k[B.a_factor], kb = B.b_factor >= 0 ? k[B.b_factor] : 1;
    VMCX_CHECK_ARG(B.w_offset >= 0, "negative w_offset");
    iv.emplace_back(B.w_offset, B.w_offset + ka * kb);
    if (B.bias_offset >= 0) iv.emplace_back(B.bias_offset, B.bias_offset + ka);
  }
  std::sort(iv.begin(), iv.end());
  int64_t at = 0;
  for (const auto &x : iv) {
    VMCX_CHECK_ARG(x.first == at, "the blocks do not tile [0, Np) (gap or overlap)");
    at = x.second;
  }
  VMCX_CHECK_ARG(at == Np, "the blocks do not tile [0, Np) (total differs from Np)");
  return 0;
}

static FactLayout fact_layout(int64_t Ns, int n_factors, const int32_t *k, int n_blocks,
                              const vmcx_sr_block *blocks) {
  FactLayout L{};
  const size_t nn = (size_t)Ns * (size_t)Ns * sizeof(double);
  int max_s = 1, max_kb = 1, max_ka = 1;
  for (int f = 0; f < n_factors; ++f) {
    L.splits[f] = gram_splits(Ns, k[f]);
    max_s = std::max(max_s, L.splits[f]);
  }
  for (int b = 0; b < n_blocks; ++b) {
    if (blocks[b].b_factor >= 0 && blocks[b].b_factor < n_factors)
      max_kb = std::max(max_kb, (int)k[blocks[b].b_factor]);
    if (blocks[b].a_factor >= 0 && blocks[b].a_factor < n_factors)
      max_ka = std::max(max_ka, (int)k[blocks[b].a_factor]);
  }
  size_t o = 0;
  L.G = o;    o += align256(nn * n_factors);
  L.Gp = o;   o += align256(nn * (max_s - 1));
  L.T = o;    o += align256(nn);
  L.Tc = o;   o += align256(nn);
  L.dinv = o; o += align256((size_t)chol_dinv_bytes(Ns));
  L.vec = o;  o += align256((size_t)Ns * 9 * sizeof(double));
  L.sc = o;   o += 256;
  L.info = o; o += 256;
  L.Y = o;    o += align256((size_t)Ns * max_kb * sizeof(double));
  L.part = o; o += align256((size_t)kColsumChunks * max_ka * sizeof(double));
  L.total = o;
  return L;
}

int64_t sr_factored_workspace_bytes(int64_t Ns, int n_factors, const int32_t *k, int n_blocks,
                                    const vmcx_sr_block *blocks) {
  if (check_factored(Ns, n_factors, k, n_blocks, blocks, -1)) return -1;
  for (int b = 0; b < n_blocks; ++b)
    if (blocks[b].b_factor >= n_factors) return -1;
  return (int64_t)fact_layout(Ns, n_factors, k, n_blocks, blocks).total;
}

// out = O^T w for a factored O (the blocks of vmcx_sr_minsr_factored), w a
// length-Ns vector: plain and bias blocks as compensated column sums of the
// factor, outer-product blocks as one FP64 GEMM each (Y: Ns x max k_b scratch,
// part: the column-sum partials).
static int factored_ot(int64_t Ns, const double *const *F, const int64_t *ld, const int32_t *k,
                       int n_blocks, const vmcx_sr_block *blocks, const double *w, double *part,
                       double *Y, double *out, rocblas_handle h, hipStream_t st) {
  const double zero = 0.0;
  const int nsi = (int)Ns;
  int rc = 0;
  for (int b = 0; b < n_blocks; ++b) {
    const vmcx_sr_block &B = blocks[b];
    const int a = B.a_factor, fb = B.b_factor;
    const double s = B.scale;
    if (fb < 0) {
      // out_block = s F_a^T w: column sums of the row-major factor
      rc = xt_gemv_cols<double>(Ns, k[a], F[a], ld[a], w, s, part, out + B.w_offset, st);
      if (rc) return rc;
      continue;
    }
    // W[i][j] = s sum_n w_n F_a[n][i] F_b[n][j]: column-major (k_b x k_a) =
    // (diag(w) F_b)^T-as-(k_b x Ns) times F_a (Ns x k_a)
    hipLaunchKernelGGL(scale_rows_kernel, dim3(2048), dim3(256), 0, st, Ns, (int)k[fb], F[fb],
                       ld[fb], w, Y);
    VMCX_LAUNCHED();
    VMCX_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, (int)k[fb],
                            (int)k[a], nsi, &s, Y, (int)k[fb], F[a], (int)ld[a], &zero,
                            out + B.w_offset, (int)k[fb]));
    if (B.bias_offset >= 0) {
      rc = xt_gemv_cols<double>(Ns, k[a], F[a], ld[a], w, s, part, out + B.bias_offset, st);
      if (rc) return rc;
    }
  }
  return 0;
}

int sr_minsr_factored(int64_t Ns, int n_factors, const double *const *F, const int64_t *ld,
                      const int32_t *k, int n_blocks, const vmcx_sr_block *blocks, int64_t Np,
                      const double *E, double energy_mean, double shift, void *ws,
                      int64_t ws_bytes, double *dp, double *stats, int32_t *info,
                      float *phase_ms, hipStream_t st) {
  VMCX_CHECK_ARG(Np >= 1, "Np must be positive");
  int rc = check_factored(Ns, n_factors, k, n_blocks, blocks, Np);
  if (rc) return rc;
  VMCX_CHECK_ARG(F && ld && E && ws && dp, "null pointer");
  for (int f = 0; f < n_factors; ++f) {
    VMCX_CHECK_ARG(F[f] != nullptr, "null factor pointer");
    VMCX_CHECK_ARG(ld[f] >= k[f] && ld[f] <= 0x7fffffff, "factor ld must be in [k, 2^31)");
  }
  const FactLayout L = fact_layout(Ns, n_factors, k, n_blocks, blocks);
  VMCX_CHECK_ARG(ws_bytes >= (int64_t)L.total, "workspace too small");
  char *w = (char *)ws;
  const int64_t nn = Ns * Ns;
  double *G = (double *)(w + L.G), *Gp = (double *)(w + L.Gp);
  double *T = (double *)(w + L.T), *Tc = (double *)(w + L.Tc), *dinv = (double *)(w + L.dinv);
  double *vec = (double *)(w + L.vec);
  double *ones = vec, *r = vec + Ns, *eps = vec + 2 * Ns, *alpha = vec + 3 * Ns;
  double *t1 = vec + 4 * Ns, *t2 = vec + 5 * Ns, *q = vec + 6 * Ns, *res = vec + 7 * Ns;
  double *at = vec + 8 * Ns;
  double *sc = (double *)(w + L.sc), *Y = (double *)(w + L.Y);
  double *part = (double *)(w + L.part);
  int32_t *iinfo = (int32_t *)(w + L.info);
  const double scale = 1.0 / std::sqrt((double)Ns), one = 1.0, zero = 0.0;
  const int nsi = (int)Ns;
  const unsigned gv = (unsigned)((Ns + 255) / 256);
  rocblas_handle h = blas_handle_for(st);
  VMCX_CHECK_ARG(h != nullptr, "rocblas_create_handle failed");
  PhaseTimer timer(phase_ms != nullptr, st);
  timer.mark();

  // factor Grams on the FP64 matrix cores
  bool used[kFactMax] = {};
  for (int b = 0; b < n_blocks; ++b) {
    used[blocks[b].a_factor] = true;
    if (blocks[b].b_factor >= 0) used[blocks[b].b_factor] = true;
  }
  for (int f = 0; f < n_factors; ++f) {
    if (!used[f]) continue;
    double *Gf = G + f * nn;
    const int sp = L.splits[f];
    rc = syrk_f64(Ns, k[f], F[f], ld[f], 1, 1.0, 0.0, Gf, Ns, sp, nn, st, Gp);
    if (rc) return rc;
    if (sp > 1) {
      hipLaunchKernelGGL(add_parts_kernel, dim3(4096), dim3(256), 0, st, Ns, sp - 1, Gp, Gf,
                         0.0);
      VMCX_LAUNCHED();
    }
  }
  timer.mark();

  // K, its double centring, eps and ||g||
  GramTerms terms{};
  terms.n = n_blocks;
  for (int b = 0; b < n_blocks; ++b) {
    terms.ga[b] = G + blocks[b].a_factor * nn;
    terms.gb[b] = blocks[b].b_factor >= 0 ? G + blocks[b].b_factor * nn : nullptr;
    terms.s2[b] = blocks[b].scale * blocks[b].scale;
    terms.bias[b] = blocks[b].bias_offset >= 0 ? 1.0 : 0.0;
  }
  hipLaunchKernelGGL(gram_combine_kernel, dim3(4096), dim3(256), 0, st, Ns, terms, T);
  hipLaunchKernelGGL(set_scalars, dim3(1), dim3(1), 0, st, sc);
  hipLaunchKernelGGL(fill_kernel, dim3(gv), dim3(256), 0, st, Ns, 1.0, ones);
  VMCX_LAUNCHED();
  VMCX_BLAS(rocblas_dsymv(h, rocblas_fill_lower, nsi, &one, T, nsi, ones, 1, &zero, r, 1));
  hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, st, Ns, r, ones, sc + 13);
  hipLaunchKernelGGL(double_center_kernel, dim3(4096), dim3(256), 0, st, Ns, r, sc + 13, shift,
                     T);
  VMCX_LAUNCHED();
  VMCX_HIP(hipMemcpyAsync(Tc, T, nn * sizeof(double), hipMemcpyDeviceToDevice, st));
  const int use_given = std::isfinite(energy_mean) ? 1 : 0;
  hipLaunchKernelGGL(energy_kernel, dim3(1), dim3(1024), 0, st, Ns, E, energy_mean, use_given,
                     scale, eps, sc + 8);
  VMCX_LAUNCHED();
  VMCX_BLAS(rocblas_dsymv(h, rocblas_fill_lower, nsi, &one, Tc, nsi, eps, 1, &zero, q, 1));
  hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, st, Ns, eps, q, sc + 14);
  hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, st, Ns, eps, eps, sc + 4);
  hipLaunchKernelGGL(gnorm_kernel, dim3(1), dim3(1), 0, st, shift, sc);
  VMCX_LAUNCHED();
  timer.mark();

  // FP64 Cholesky (the forward half of the solve runs along) and solve
  VMCX_HIP(hipMemcpyAsync(t1, eps, Ns * sizeof(double), hipMemcpyDeviceToDevice, st));
  rc = chol_factor((void *)h, Ns, T, Ns, dinv, iinfo, st, phase_ms ? phase_ms + 5 : nullptr,
                   t1, t2);
  if (rc) return rc;
  timer.mark();
  rc = chol_solve_bwd(Ns, chol_panels(Ns, dinv), Ns, dinv, t2, alpha, st);
  if (rc) return rc;
  timer.mark();

  if (stats) {
    // residual of the solve against the FP64 Gram
    VMCX_BLAS(rocblas_dsymv(h, rocblas_fill_lower, nsi, &one, Tc, nsi, alpha, 1, &zero, q, 1));
    hipLaunchKernelGGL(residual2_kernel, dim3(gv), dim3(256), 0, st, Ns, eps, q, alpha, 0.0,
                       res);
    hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, st, Ns, res, sc + 10);
    hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, st, Ns, eps, sc + 11);
    VMCX_LAUNCHED();
  }
  // dp = O^T a~, a~ = (alpha - mean alpha) / sqrt(Ns)
  hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, st, Ns, alpha, ones, sc + 15);
  hipLaunchKernelGGL(center_vec_kernel, dim3(gv), dim3(256), 0, st, Ns, alpha, sc + 15, scale,
                     at);
  VMCX_LAUNCHED();
  rc = factored_ot(Ns, F, ld, k, n_blocks, blocks, at, part, Y, dp, h, st);
  if (rc) return rc;
  if (stats) {
    hipLaunchKernelGGL(finish_stats, dim3(1), dim3(1), 0, st, sc, stats);
    VMCX_LAUNCHED();
  }
  if (info) VMCX_HIP(hipMemcpyAsync(info, iinfo, sizeof(int32_t), hipMemcpyDeviceToDevice, st));
  timer.mark();
  timer.read(phase_ms);
  return 0;
}

// The energy gradient g = O^T (E - E_mean) / Ns on a factored O (the
// reference's run_SR = False branch): no Gram, no factorisation.
int sr_factored_gradient(int64_t Ns, int n_factors, const double *const *F, const int64_t *ld,
                         const int32_t *k, int n_blocks, const vmcx_sr_block *blocks, int64_t Np,
                         const double *E, double energy_mean, void *ws, int64_t ws_bytes,
                         double *g, double *stats, hipStream_t st) {
  VMCX_CHECK_ARG(Np >= 1, "Np must be positive");
  int rc = check_factored(Ns, n_factors, k, n_blocks, blocks, Np);
  if (rc) return rc;
  VMCX_CHECK_ARG(F && ld && E && ws && g, "null pointer");
  for (int f = 0; f < n_factors; ++f) {
    VMCX_CHECK_ARG(F[f] != nullptr, "null factor pointer");
    VMCX_CHECK_ARG(ld[f] >= k[f] && ld[f] <= 0x7fffffff, "factor ld must be in [k, 2^31)");
  }
  const FactLayout L = fact_layout(Ns, n_factors, k, n_blocks, blocks);
  VMCX_CHECK_ARG(ws_bytes >= (int64_t)L.total, "workspace too small");
  char *w = (char *)ws;
  double *vec = (double *)(w + L.vec), *sc = (double *)(w + L.sc);
  double *Y = (double *)(w + L.Y), *part = (double *)(w + L.part);
  double *wE = vec;
  rocblas_handle h = blas_handle_for(st);
  VMCX_CHECK_ARG(h != nullptr, "rocblas_create_handle failed");
  hipLaunchKernelGGL(set_scalars, dim3(1), dim3(1), 0, st, sc);
  const int use_given = std::isfinite(energy_mean) ? 1 : 0;
  // wE = (E - E_mean) / Ns (compensated mean and variance into sc[8], sc[9])
  hipLaunchKernelGGL(energy_kernel, dim3(1), dim3(1024), 0, st, Ns, E, energy_mean, use_given,
                     1.0 / (double)Ns, wE, sc + 8);
  VMCX_LAUNCHED();
  rc = factored_ot(Ns, F, ld, k, n_blocks, blocks, wE, part, Y, g, h, st);
  if (rc) return rc;
  if (stats) {
    hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, st, Np, g, sc + 12);
    hipLaunchKernelGGL(finish_stats, dim3(1), dim3(1), 0, st, sc, stats);
    VMCX_LAUNCHED();
  }
  return 0;
}

// [mean, variance] of E: compensated sum for the mean, then a compensated
// sum of (E - mean)^2 (one workgroup; the driver's energy statistics).
__global__ void __launch_bounds__(1024)
    energy_stats_kernel(int64_t n, const double *__restrict__ E, double *out) {
  __shared__ double sh[2][16];
  __shared__ double mean_sh;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int pass = 0; pass < 2; ++pass) {
    const double m = pass ? mean_sh : 0.0;
    Neu a;
    for (int64_t s = threadIdx.x; s < n; s += blockDim.x) {
      const double v = pass ? (E[s] - m) * (E[s] - m) : E[s];
      a.add(v);
    }
    double v0 = a.s, v1 = a.c;
    for (int o = 32; o > 0; o >>= 1) {
      Neu x;
      x.s = v0; x.c = v1; x.add(__shfl_down(v0, o, 64)); x.c += __shfl_down(v1, o, 64);
      v0 = x.s; v1 = x.c;
    }
    if (lane == 0) {
      sh[0][wv] = v0;
      sh[1][wv] = v1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      Neu x;
      for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
        x.add(sh[0][k]);
        x.c += sh[1][k];
      }
      const double r = x.value() / (double)n;
      if (pass == 0) {
        mean_sh = r;
        out[0] = r;
      } else {
        out[1] = r;
      }
    }
    __syncthreads();
  }
}

int energy_stats(int64_t n, const double *E, double *out, hipStream_t st) {
  VMCX_CHECK_ARG(n >= 1, "n must be positive");
  VMCX_CHECK_ARG(E && out, "null pointer");
  hipLaunchKernelGGL(energy_stats_kernel, dim3(1), dim3(1024), 0, st, n, E, out);
  VMCX_LAUNCHED();
  return 0;
}

// p <- p - lr * dp
int sgd_update(int64_t n, double *params, const double *dp, double lr, hipStream_t st) {
  VMCX_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return 0;
  VMCX_CHECK_ARG(params && dp, "null pointer");
  hipLaunchKernelGGL(axpy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                     -lr, dp, params);
  VMCX_LAUNCHED();
  return 0;
}

}  // namespace vmcx
