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
// Gram, GEMV and Cholesky are plain library BLAS / LAPACK calls (rocBLAS,
// rocSOLVER); the fused centring / conversion / shift kernels are ours.
#include <hip/hip_bf16.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <cstdlib>
#include <mutex>
#include <unordered_map>

#include "kernels.h"

namespace vmcx {

// ------------------------------------------------------------ reductions
struct Neu {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

// partial[r][p] = sum over rows in chunk r of O[s][p]   (grid: (ceil(Np/256), R))
__global__ void __launch_bounds__(256)
    colsum_partial_kernel(int64_t Ns, int64_t Np, const double *__restrict__ O,
                          int64_t ldo, int64_t rows_per_chunk, double *partial) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
  const int64_t r1 = (r0 + rows_per_chunk < Ns) ? r0 + rows_per_chunk : Ns;
  Neu acc;
  for (int64_t s = r0; s < r1; ++s) acc.add(O[s * ldo + p]);
  partial[(int64_t)blockIdx.y * Np + p] = acc.value();
}

__global__ void __launch_bounds__(256)
    colsum_final_kernel(int64_t Np, int R, const double *__restrict__ partial,
                        double inv_n, double *mean) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  Neu acc;
  for (int r = 0; r < R; ++r) acc.add(partial[(int64_t)r * Np + p]);
  mean[p] = acc.value() * inv_n;
}

// X[s][p] = (O[s][p] - mean[p]) * scale ; optional low-precision copy
template <typename TL>
__global__ void __launch_bounds__(256)
    center_kernel(int64_t Ns, int64_t Np, const double *__restrict__ O,
                  int64_t ldo, const double *__restrict__ mean, double scale,
                  double *X, TL *Xlo) {
  const int64_t n = Ns * Np;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = idx / Np, p = idx - s * Np;
    const double v = (O[s * ldo + p] - mean[p]) * scale;
    X[idx] = v;
    if (Xlo) Xlo[idx] = (TL)v;
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
rocblas_handle blas_handle_for(hipStream_t st) {
  static std::mutex mu;
  static std::unordered_map<int, rocblas_handle> handles;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  auto it = handles.find(dev);
  rocblas_handle h;
  if (it == handles.end()) {
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    handles[dev] = h;
  } else {
    h = it->second;
  }
  rocblas_set_stream(h, st);
  rocblas_set_pointer_mode(h, rocblas_pointer_mode_host);
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
  size_t X, Xlo, T, Tlo, partial, mean, eps, alpha, r, u, v, stats, info, dinv, t1, t2,
      corr, total;
};

static SrLayout sr_layout(int64_t Ns, int64_t Np, int gram) {
  SrLayout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  const int R = 64;
  L.X = take((size_t)Ns * Np * 8);
  L.Xlo = take(gram == 1 ? (size_t)Ns * Np * 4 : (gram == 2 ? (size_t)Ns * Np * 2 : 0));
  L.T = take((size_t)Ns * Ns * 8);
  L.Tlo = take(gram ? (size_t)Ns * Ns * 4 : 0);
  L.partial = take((size_t)R * Np * 8);
  L.mean = take((size_t)Np * 8);
  L.eps = take((size_t)Ns * 8);
  L.alpha = take((size_t)Ns * 8);
  L.r = take((size_t)Ns * 8);
  L.u = take((size_t)Np * 8);
  L.v = take((size_t)Ns * 8);
  L.stats = take(8 * 8);
  L.info = take(64);
  L.dinv = take((size_t)chol_dinv_bytes(Ns));
  L.t1 = take((size_t)Ns * 8);
  L.t2 = take((size_t)Ns * 8);
  L.corr = take((size_t)Ns * 8);
  L.total = off;
  return L;
}

int64_t sr_workspace_bytes(int64_t Ns, int64_t Np, int gram) {
  if (Ns <= 0 || Np <= 0 || gram < 0 || gram > 2) return -1;
  return (int64_t)sr_layout(Ns, Np, gram).total;
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

int sr_minsr(int64_t Ns, int64_t Np, const double *O, int64_t ldo,
             const double *E, double energy_mean, double shift, int gram,
             int refine, void *ws, int64_t ws_bytes, double *dp, double *stats,
             int32_t *info, float *phase_ms, hipStream_t st) {
  VMCX_CHECK_ARG(Ns >= 1 && Np >= 1, "Ns and Np must be positive");
  VMCX_CHECK_ARG(Ns <= 0x7fffffff && Np <= 0x7fffffff, "Ns / Np exceed 32-bit BLAS");
  VMCX_CHECK_ARG(ldo >= Np, "ldo < Np");
  VMCX_CHECK_ARG(gram >= 0 && gram <= 2, "gram must be 0 (f64), 1 (f32) or 2 (bf16)");
  VMCX_CHECK_ARG(refine >= 0 && refine <= 32, "refine must be in [0, 32]");
  VMCX_CHECK_ARG(O && E && ws && dp, "null pointer");
  const SrLayout L = sr_layout(Ns, Np, gram);
  VMCX_CHECK_ARG(ws_bytes >= (int64_t)L.total, "workspace too small");
  char *w = (char *)ws;
  double *X = (double *)(w + L.X);
  double *T = (double *)(w + L.T);
  double *partial = (double *)(w + L.partial), *mean = (double *)(w + L.mean);
  double *eps = (double *)(w + L.eps), *alpha = (double *)(w + L.alpha);
  double *r = (double *)(w + L.r), *u = (double *)(w + L.u), *v = (double *)(w + L.v);
  double *stw = (double *)(w + L.stats);
  rocblas_int *dinfo = (rocblas_int *)(w + L.info);
  const double scale = 1.0 / std::sqrt((double)Ns);
  const int nsi = (int)Ns, npi = (int)Np;

  rocblas_handle h = blas_handle_for(st);
  VMCX_CHECK_ARG(h != nullptr, "rocblas_create_handle failed");
  PhaseTimer timer(phase_ms != nullptr, st);
  timer.mark();

  // column means (two-level compensated sum) and centring
  const int R = 64;
  const int64_t rows = (Ns + R - 1) / R;
  hipLaunchKernelGGL(colsum_partial_kernel, dim3((unsigned)((Np + 255) / 256), R),
                     dim3(256), 0, st, Ns, Np, O, ldo, rows, partial);
  VMCX_LAUNCHED();
  hipLaunchKernelGGL(colsum_final_kernel, dim3((unsigned)((Np + 255) / 256)), dim3(256),
                     0, st, Np, R, partial, 1.0 / (double)Ns, mean);
  VMCX_LAUNCHED();
  const unsigned cgrid = 4096;
  if (gram == 1)
    hipLaunchKernelGGL(center_kernel<float>, dim3(cgrid), dim3(256), 0, st, Ns, Np, O,
                       ldo, mean, scale, X, (float *)(w + L.Xlo));
  else if (gram == 2)
    hipLaunchKernelGGL(center_kernel<__hip_bfloat16>, dim3(cgrid), dim3(256), 0, st, Ns,
                       Np, O, ldo, mean, scale, X, (__hip_bfloat16 *)(w + L.Xlo));
  else
    hipLaunchKernelGGL(center_kernel<float>, dim3(cgrid), dim3(256), 0, st, Ns, Np, O,
                       ldo, mean, scale, X, (float *)nullptr);
  VMCX_LAUNCHED();
  const int use_given = std::isfinite(energy_mean) ? 1 : 0;
  hipLaunchKernelGGL(energy_kernel, dim3(1), dim3(1024), 0, st, Ns, E, energy_mean,
                     use_given, scale, eps, stw);
  VMCX_LAUNCHED();

  timer.mark();
  // Gram
  const double one = 1.0, zero = 0.0;
  if (gram == 0) {
    VMCX_BLAS(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, nsi, npi,
                            &one, X, npi, &zero, T, nsi));
    hipLaunchKernelGGL(diag_shift_kernel, dim3((unsigned)((Ns + 255) / 256)), dim3(256), 0,
                       st, Ns, T, shift);
    VMCX_LAUNCHED();
  } else {
    const float onef = 1.0f, zerof = 0.0f;
    const rocblas_datatype ab = (gram == 1) ? rocblas_datatype_f32_r : rocblas_datatype_bf16_r;
    void *Xlo = (void *)(w + L.Xlo);
    float *Tlo = (float *)(w + L.Tlo);
    VMCX_BLAS(rocblas_gemm_ex(h, rocblas_operation_transpose, rocblas_operation_none, nsi,
                              nsi, npi, &onef, Xlo, ab, npi, Xlo, ab, npi, &zerof, Tlo,
                              rocblas_datatype_f32_r, nsi, Tlo, rocblas_datatype_f32_r,
                              nsi, rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0,
                              0));
    hipLaunchKernelGGL(shift_kernel<float>, dim3(4096), dim3(256), 0, st, Ns, Tlo, T,
                       shift);
    VMCX_LAUNCHED();
  }
  // FP64 Cholesky and solve (blocked factorisation of chol.hip; the
  // rocSOLVER path is kept for cross-checks: VMCX_SR_ROCSOLVER=1)
  static const bool use_rocsolver = getenv("VMCX_SR_ROCSOLVER") != nullptr;
  double *t1 = (double *)(w + L.t1), *t2 = (double *)(w + L.t2);
  double *corr = (double *)(w + L.corr), *dinv = (double *)(w + L.dinv);
  int32_t *iinfo = (int32_t *)dinfo;
  timer.mark();
  if (use_rocsolver) {
    VMCX_BLAS(rocsolver_dpotrf(h, rocblas_fill_lower, nsi, T, nsi, dinfo));
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((Ns + 255) / 256)), dim3(256), 0, st,
                       Ns, eps, alpha);
    VMCX_LAUNCHED();
    VMCX_BLAS(rocsolver_dpotrs(h, rocblas_fill_lower, nsi, 1, T, nsi, alpha, nsi));
  } else {
    int rc = chol_factor((void *)h, Ns, T, Ns, dinv, iinfo, st);
    if (rc) return rc;
    timer.mark();
    rc = chol_solve(Ns, T, Ns, dinv, eps, alpha, t1, t2, st);
    if (rc) return rc;
  }
  if (use_rocsolver) timer.mark();
  // iterative refinement against the FP64 operator  Ot Ot^T + shift I
  for (int it = 0; it < refine; ++it) {
    VMCX_BLAS(rocblas_dgemv(h, rocblas_operation_none, npi, nsi, &one, X, npi, alpha, 1,
                            &zero, u, 1));
    VMCX_BLAS(rocblas_dgemv(h, rocblas_operation_transpose, npi, nsi, &one, X, npi, u, 1,
                            &zero, v, 1));
    hipLaunchKernelGGL(residual_kernel, dim3((unsigned)((Ns + 255) / 256)), dim3(256), 0,
                       st, Ns, eps, v, alpha, shift, r);
    VMCX_LAUNCHED();
    if (use_rocsolver) {
      VMCX_BLAS(rocsolver_dpotrs(h, rocblas_fill_lower, nsi, 1, T, nsi, r, nsi));
      hipLaunchKernelGGL(axpy_kernel, dim3((unsigned)((Ns + 255) / 256)), dim3(256), 0, st,
                         Ns, 1.0, r, alpha);
    } else {
      const int rc = chol_solve(Ns, T, Ns, dinv, r, corr, t1, t2, st);
      if (rc) return rc;
      hipLaunchKernelGGL(axpy_kernel, dim3((unsigned)((Ns + 255) / 256)), dim3(256), 0, st,
                         Ns, 1.0, corr, alpha);
    }
    VMCX_LAUNCHED();
  }
  timer.mark();
  // dp = Ot^T alpha
  VMCX_BLAS(rocblas_dgemv(h, rocblas_operation_none, npi, nsi, &one, X, npi, alpha, 1, &zero,
                          dp, 1));
  if (stats) {
    // final residual norm of (Ot Ot^T + shift) alpha - eps and the gradient norm
    VMCX_BLAS(rocblas_dgemv(h, rocblas_operation_transpose, npi, nsi, &one, X, npi, dp, 1,
                            &zero, v, 1));
    hipLaunchKernelGGL(residual_kernel, dim3((unsigned)((Ns + 255) / 256)), dim3(256), 0,
                       st, Ns, eps, v, alpha, shift, r);
    VMCX_LAUNCHED();
    hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, st, Ns, r, stw + 2);
    VMCX_LAUNCHED();
    VMCX_BLAS(rocblas_dgemv(h, rocblas_operation_none, npi, nsi, &one, X, npi, eps, 1, &zero,
                            u, 1));
    hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, st, Np, u, stw + 3);
    VMCX_LAUNCHED();
    VMCX_HIP(hipMemcpyAsync(stats, stw, 4 * sizeof(double), hipMemcpyDeviceToDevice, st));
  }
  if (info) VMCX_HIP(hipMemcpyAsync(info, dinfo, sizeof(int32_t), hipMemcpyDeviceToDevice, st));
  timer.mark();
  timer.read(phase_ms);
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
