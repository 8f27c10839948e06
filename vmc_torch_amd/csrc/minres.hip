// Parameter-space stochastic reconfiguration by MINRES, resident on the
// device: (S + shift I) dp = g with
//   S x = O^T (O x) / Ns - (Obar . x) Obar,   Obar = mean_s O,
//   g   = (1/Ns) sum_s (E_s - E_mean) O_s,
// S never formed.  The Lanczos / Givens recurrence of Paige and Saunders
// (the form of scipy.sparse.linalg.minres: convergence when
// |phibar| < rtol ||A|| ||x|| with the running estimate of ||A||) keeps all
// of its scalars in device memory; one-thread kernels update them, and a
// device flag set on convergence turns every later launch of the loop into
// an immediate return, so the host never waits inside the solve.
//
// Per iteration: one pass over O for O v, one for O^T (O v) (FP64
// accumulation for either O dtype), five length-Np reductions.
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.h"

namespace vmcx {

namespace {

struct Neu2 {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
};

constexpr int kRowChunks = 64;     // row chunks of the O^T w column sums
constexpr int kDotBlocks = 1024;   // first-stage blocks of a length-Np dot

// scalar slots
enum {
  S_BETA = 0, S_OLDB, S_ALFA, S_CS, S_SN, S_DBAR, S_EPSLN, S_PHIBAR, S_ANORM2, S_OLDEPS,
  S_DELTA, S_GAMMA, S_PHI, S_DONE, S_ITN, S_MV, S_YNORM2, S_BETA2, S_BNORM2, S_INFO,
  S_EMEAN, S_EVAR, S_ZERO = 31, S_COUNT = 32
};

__global__ void __launch_bounds__(256)
    mr_colsum_partial(int64_t Ns, int64_t Np, const void *O, int o_dtype, int64_t ldo,
                      int64_t rows, const double *__restrict__ w, const double *__restrict__ w2,
                      double *partial, double *partial2, const double *skip) {
  if (skip[0] != 0.0) return;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  const int64_t r0 = (int64_t)blockIdx.y * rows;
  const int64_t r1 = (r0 + rows < Ns) ? r0 + rows : Ns;
  Neu2 a, b;
  for (int64_t s = r0; s < r1; ++s) {
    const double v = o_dtype == 0 ? ((const double *)O)[s * ldo + p]
                                  : (double)((const float *)O)[s * ldo + p];
    a.add(w ? v * w[s] : v);
    if (w2) b.add(v * w2[s]);
  }
  partial[(int64_t)blockIdx.y * Np + p] = a.s + a.c;
  if (w2) partial2[(int64_t)blockIdx.y * Np + p] = b.s + b.c;
}

__global__ void __launch_bounds__(256)
    mr_colsum_final(int64_t Np, const double *partial, double scale, double *out,
                    const double *skip) {
  if (skip[0] != 0.0) return;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  Neu2 a;
  for (int r = 0; r < kRowChunks; ++r) a.add(partial[(int64_t)r * Np + p]);
  out[p] = (a.s + a.c) * scale;
}

// v[s] = sum_p O[s][p] x[p]
__global__ void __launch_bounds__(256)
    mr_gemv_rows(int64_t Np, const void *O, int o_dtype, int64_t ldo, const double *x,
                 double *v, const double *sc) {
  if (sc[S_DONE] != 0.0) return;
  __shared__ double sh[4];
  const int64_t s = blockIdx.x;
  double acc = 0.0;
  if (o_dtype == 0) {
    const double *row = (const double *)O + s * ldo;
    for (int64_t p = threadIdx.x; p < Np; p += blockDim.x) acc += row[p] * x[p];
  } else {
    const float *row = (const float *)O + s * ldo;
    for (int64_t p = threadIdx.x; p < Np; p += blockDim.x) acc += (double)row[p] * x[p];
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) v[s] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// two-stage dot product a . b -> sc[slot]
__global__ void __launch_bounds__(256)
    mr_dot_partial(int64_t n, const double *a, const double *b, double *part,
                   const double *skip) {
  if (skip[0] != 0.0) return;
  __shared__ double sh[4];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    acc += a[i] * b[i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ void __launch_bounds__(1024)
    mr_dot_final(int nb, const double *part, const double *skip, double *sc, int slot) {
  if (skip[0] != 0.0) return;
  __shared__ double sh[16];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += part[i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    Neu2 a;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) a.add(sh[k]);
    sc[slot] = a.s + a.c;
  }
}

// E statistics and w = E - E_mean (one block)
__global__ void __launch_bounds__(1024)
    mr_energy(int64_t Ns, const double *E, double given_mean, int use_given, double *w,
              double *sc) {
  __shared__ double sh[2][16];
  __shared__ double mean_sh;
  Neu2 a, b;
  for (int64_t s = threadIdx.x; s < Ns; s += blockDim.x) {
    a.add(E[s]);
    b.add(E[s] * E[s]);
  }
  double v0 = a.s + a.c, v1 = b.s + b.c;
  for (int o = 32; o > 0; o >>= 1) {
    v0 += __shfl_down(v0, o, 64);
    v1 += __shfl_down(v1, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = v0;
    sh[1][threadIdx.x >> 6] = v1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Neu2 x, y;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
      x.add(sh[0][k]);
      y.add(sh[1][k]);
    }
    const double m = (x.s + x.c) / (double)Ns;
    sc[S_EMEAN] = m;
    sc[S_EVAR] = (y.s + y.c) / (double)Ns - m * m;
    mean_sh = use_given ? given_mean : m;
  }
  __syncthreads();
  const double em = mean_sh;
  for (int64_t s = threadIdx.x; s < Ns; s += blockDim.x) w[s] = E[s] - em;
}

__global__ void mr_zero(int64_t n, double *x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

__global__ void mr_copy(int64_t n, const double *a, double *b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// beta1 = ||g||; trivial right-hand side ends the solve with x = 0
__global__ void mr_init(double *sc, int maxiter) {
  for (int k = 0; k < S_COUNT; ++k)
    if (k != S_EMEAN && k != S_EVAR && k != S_BNORM2) sc[k] = 0.0;
  const double bn = sqrt(sc[S_BNORM2]);
  sc[S_BETA] = bn;
  sc[S_CS] = -1.0;
  sc[S_PHIBAR] = bn;
  sc[S_INFO] = (double)maxiter;
  if (bn == 0.0) {
    sc[S_DONE] = 1.0;
    sc[S_INFO] = 0.0;
  }
}

// v = r2 / beta
__global__ void mr_scale_v(int64_t n, const double *r2, double *v, const double *sc) {
  if (sc[S_DONE] != 0.0) return;
  const double s = 1.0 / sc[S_BETA];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = s * r2[i];
}

// y = Sx - (Obar . v) Obar + shift v - (beta / oldb) r1   (last term after
// the first iteration)
__global__ void mr_form_y(int64_t n, const double *Sx, const double *m, const double *v,
                          const double *r1, double shift, double *y, const double *sc) {
  if (sc[S_DONE] != 0.0) return;
  const double mv = sc[S_MV];
  const double c = sc[S_ITN] >= 1.0 ? sc[S_BETA] / sc[S_OLDB] : 0.0;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double t = Sx[i] - mv * m[i] + shift * v[i];
    if (c != 0.0) t -= c * r1[i];
    y[i] = t;
  }
}

// y -= (alfa / beta) r2
__global__ void mr_orth_y(int64_t n, const double *r2, double *y, const double *sc) {
  if (sc[S_DONE] != 0.0) return;
  const double c = sc[S_ALFA] / sc[S_BETA];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] -= c * r2[i];
}

// Givens rotation of the new Lanczos column (sc[S_BETA2] = ||y||^2)
__global__ void mr_rotate(double *sc) {
  if (sc[S_DONE] != 0.0) return;
  const double alfa = sc[S_ALFA];
  const double oldb = sc[S_BETA];
  const double beta = sqrt(sc[S_BETA2]);
  const double cs = sc[S_CS], sn = sc[S_SN], dbar = sc[S_DBAR];
  sc[S_OLDB] = oldb;
  sc[S_BETA] = beta;
  sc[S_OLDEPS] = sc[S_EPSLN];
  sc[S_DELTA] = cs * dbar + sn * alfa;
  const double gbar = sn * dbar - cs * alfa;
  sc[S_EPSLN] = sn * beta;
  sc[S_DBAR] = -cs * beta;
  double gamma = sqrt(gbar * gbar + beta * beta);
  gamma = gamma > 1e-300 ? gamma : 1e-300;
  sc[S_GAMMA] = gamma;
  const double ncs = gbar / gamma, nsn = beta / gamma;
  sc[S_CS] = ncs;
  sc[S_SN] = nsn;
  sc[S_PHI] = ncs * sc[S_PHIBAR];
  sc[S_PHIBAR] = nsn * sc[S_PHIBAR];
  sc[S_ANORM2] += alfa * alfa + oldb * oldb + beta * beta;
}

// w = (v - oldeps w1 - delta w2) / gamma ; x += phi w
__global__ void mr_update_x(int64_t n, const double *v, const double *w1, const double *w2,
                            double *w, double *x, const double *sc) {
  if (sc[S_DONE] != 0.0) return;
  const double oe = sc[S_OLDEPS], de = sc[S_DELTA], ig = 1.0 / sc[S_GAMMA], ph = sc[S_PHI];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double wi = (v[i] - oe * w1[i] - de * w2[i]) * ig;
    w[i] = wi;
    x[i] += ph * wi;
  }
}

__global__ void mr_check(double rtol, double *sc) {
  if (sc[S_DONE] != 0.0) return;
  sc[S_ITN] += 1.0;
  const double rnorm = fabs(sc[S_PHIBAR]);
  if (rnorm < rtol * sqrt(sc[S_ANORM2]) * sqrt(sc[S_YNORM2]) || sc[S_BETA] == 0.0) {
    sc[S_DONE] = 1.0;
    sc[S_INFO] = 0.0;
  }
}

// stats <- [mean E, var E, |phibar| (residual estimate), ||g||, iterations]
__global__ void mr_stats(const double *sc, double *stats, int32_t *info) {
  if (stats) {
    stats[0] = sc[S_EMEAN];
    stats[1] = sc[S_EVAR];
    stats[2] = fabs(sc[S_PHIBAR]);
    stats[3] = sqrt(sc[S_BNORM2]);
    stats[4] = sc[S_ITN];
  }
  if (info) info[0] = (int32_t)sc[S_INFO];
}

inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

struct MrLayout {
  size_t vec[9], m, g, w, inner, partial, dotp, sc, total;
};

MrLayout mr_layout(int64_t Ns, int64_t Np) {
  MrLayout L{};
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t o = off;
    off += al256(b);
    return o;
  };
  for (auto &v : L.vec) v = take((size_t)Np * 8);
  L.m = take((size_t)Np * 8);
  L.g = take((size_t)Np * 8);
  L.w = take((size_t)Ns * 8);
  L.inner = take((size_t)Ns * 8);
  L.partial = take((size_t)2 * kRowChunks * Np * 8);
  L.dotp = take((size_t)kDotBlocks * 8);
  L.sc = take((size_t)S_COUNT * 8);
  L.total = off;
  return L;
}

}  // namespace

int64_t sr_minres_workspace_bytes(int64_t Ns, int64_t Np) {
  if (Ns <= 0 || Np <= 0) return -1;
  return (int64_t)mr_layout(Ns, Np).total;
}

int sr_minres(int64_t Ns, int64_t Np, const void *O, int o_dtype, int64_t ldo, const double *E,
              double energy_mean, double diag_shift, double rtol, int maxiter, void *ws,
              int64_t ws_bytes, double *dp, double *stats, int32_t *info, hipStream_t st) {
  VMCX_CHECK_ARG(Ns >= 1 && Np >= 1, "Ns and Np must be positive");
  VMCX_CHECK_ARG(ldo >= Np, "ldo < Np");
  VMCX_CHECK_ARG(o_dtype == 0 || o_dtype == 1, "o_dtype must be 0 (f64) or 1 (f32)");
  VMCX_CHECK_ARG(maxiter >= 0, "maxiter must be >= 0");
  VMCX_CHECK_ARG(rtol >= 0.0, "rtol must be >= 0");
  VMCX_CHECK_ARG(O && E && dp && ws, "null pointer argument");
  const MrLayout L = mr_layout(Ns, Np);
  VMCX_CHECK_ARG(ws_bytes >= (int64_t)L.total, "workspace too small");
  char *w = (char *)ws;
  double *vecs[9];
  for (int k = 0; k < 9; ++k) vecs[k] = (double *)(w + L.vec[k]);
  double *m = (double *)(w + L.m), *g = (double *)(w + L.g), *wE = (double *)(w + L.w);
  double *inner = (double *)(w + L.inner), *partial = (double *)(w + L.partial);
  double *partial2 = partial + (int64_t)kRowChunks * Np;
  double *dotp = (double *)(w + L.dotp), *sc = (double *)(w + L.sc);
  double *x = dp;
  const unsigned gn = (unsigned)((Np + 255) / 256);
  const int64_t rows = (Ns + kRowChunks - 1) / kRowChunks;
  const int nb = (int)((Np + 2047) / 2048 < kDotBlocks ? (Np + 2047) / 2048 : kDotBlocks);
  const int nbd = nb > 0 ? nb : 1;
  // launches outside the loop take the flag that stays zero, those inside the
  // convergence flag
  const double *always = sc + S_ZERO, *done = sc + S_DONE;
  auto dot = [&](const double *a, const double *b, int slot, const double *skip) {
    hipLaunchKernelGGL(mr_dot_partial, dim3(nbd), dim3(256), 0, st, Np, a, b, dotp, skip);
    hipLaunchKernelGGL(mr_dot_final, dim3(1), dim3(1024), 0, st, nbd, dotp, skip, sc, slot);
  };
  hipLaunchKernelGGL(mr_zero, dim3(1), dim3(S_COUNT), 0, st, (int64_t)S_COUNT, sc);
  VMCX_LAUNCHED();
  // E statistics, w = E - E_mean; Obar and g from one read of O
  const int use_given = std::isfinite(energy_mean) ? 1 : 0;
  hipLaunchKernelGGL(mr_energy, dim3(1), dim3(1024), 0, st, Ns, E, energy_mean, use_given, wE,
                     sc);
  hipLaunchKernelGGL(mr_colsum_partial, dim3(gn, kRowChunks), dim3(256), 0, st, Ns, Np, O,
                     o_dtype, ldo, rows, (const double *)nullptr, (const double *)wE, partial,
                     partial2, always);
  hipLaunchKernelGGL(mr_colsum_final, dim3(gn), dim3(256), 0, st, Np, partial,
                     1.0 / (double)Ns, m, always);
  hipLaunchKernelGGL(mr_colsum_final, dim3(gn), dim3(256), 0, st, Np, partial2,
                     1.0 / (double)Ns, g, always);
  VMCX_LAUNCHED();
  dot(g, g, S_BNORM2, always);
  if (maxiter == 0)  // the energy gradient only
    hipLaunchKernelGGL(mr_copy, dim3(gn), dim3(256), 0, st, Np, g, x);
  else
    hipLaunchKernelGGL(mr_zero, dim3(gn), dim3(256), 0, st, Np, x);
  hipLaunchKernelGGL(mr_init, dim3(1), dim3(1), 0, st, sc, maxiter);
  // Lanczos vectors: r1, r2, y rotate through vecs[0..2]; w1, w2, w through
  // vecs[3..5]; v = vecs[6]; Sx = vecs[7]
  double *r1 = vecs[0], *r2 = vecs[1], *y = vecs[2];
  double *w1 = vecs[3], *w2 = vecs[4], *wn = vecs[5];
  double *v = vecs[6], *Sx = vecs[7];
  hipLaunchKernelGGL(mr_copy, dim3(gn), dim3(256), 0, st, Np, g, r1);
  hipLaunchKernelGGL(mr_copy, dim3(gn), dim3(256), 0, st, Np, g, r2);
  hipLaunchKernelGGL(mr_zero, dim3(gn), dim3(256), 0, st, Np, w1);
  hipLaunchKernelGGL(mr_zero, dim3(gn), dim3(256), 0, st, Np, w2);
  hipLaunchKernelGGL(mr_zero, dim3(gn), dim3(256), 0, st, Np, wn);
  VMCX_LAUNCHED();
  for (int it = 0; it < maxiter; ++it) {
    hipLaunchKernelGGL(mr_scale_v, dim3(gn), dim3(256), 0, st, Np, r2, v, sc);
    // S v without S: O v, then O^T (O v) / Ns
    hipLaunchKernelGGL(mr_gemv_rows, dim3((unsigned)Ns), dim3(256), 0, st, Np, O, o_dtype, ldo,
                       v, inner, sc);
    hipLaunchKernelGGL(mr_colsum_partial, dim3(gn, kRowChunks), dim3(256), 0, st, Ns, Np, O,
                       o_dtype, ldo, rows, (const double *)inner, (const double *)nullptr,
                       partial, partial2, done);
    hipLaunchKernelGGL(mr_colsum_final, dim3(gn), dim3(256), 0, st, Np, partial,
                       1.0 / (double)Ns, Sx, done);
    dot(m, v, S_MV, done);
    hipLaunchKernelGGL(mr_form_y, dim3(gn), dim3(256), 0, st, Np, Sx, m, v, r1, diag_shift, y,
                       sc);
    dot(v, y, S_ALFA, done);
    hipLaunchKernelGGL(mr_orth_y, dim3(gn), dim3(256), 0, st, Np, r2, y, sc);
    dot(y, y, S_BETA2, done);
    hipLaunchKernelGGL(mr_rotate, dim3(1), dim3(1), 0, st, sc);
    hipLaunchKernelGGL(mr_update_x, dim3(gn), dim3(256), 0, st, Np, v, w1, w2, wn, x, sc);
    dot(x, x, S_YNORM2, done);
    hipLaunchKernelGGL(mr_check, dim3(1), dim3(1), 0, st, rtol, sc);
    VMCX_LAUNCHED();
    // r1 <- r2 <- y ; w1 <- w2 <- w (buffers rotate on the host; once the
    // device flag is set nothing reads them again)
    double *t = r1;
    r1 = r2;
    r2 = y;
    y = t;
    t = w1;
    w1 = w2;
    w2 = wn;
    wn = t;
  }
  hipLaunchKernelGGL(mr_stats, dim3(1), dim3(1), 0, st, sc, stats, info);
  VMCX_LAUNCHED();
  return 0;
}

}  // namespace vmcx
