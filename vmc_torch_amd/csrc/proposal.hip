// Learned Metropolis proposal on gfx950: a 3-layer 3x3 convolutional net on
// the (n_up, n_dn) occupation grid gives logits for every local move
// m = 3 e + t (edges e = (i, j), i < j, in sorted order; t = hop of an up
// electron, hop of a down electron, exchange of a down and an up electron);
// q(m | x) is the softmax over the valid moves and every move is its own
// reverse, so the Metropolis-Hastings ratio
//   |psi(x')|^2 / |psi(x)|^2 * q(m | x') / q(m | x)
// needs the net at x and at x' (two launches of the same kernel).
//
// One workgroup per configuration: activations [C][N] in LDS (C <= 32,
// N <= 64), FP32 convolutions with zero padding; logits and log-sum-exp in
// FP64.  The backward pass (for the heat-bath training target) recomputes
// the pre-activations and writes one gradient row per configuration, summed
// over configurations by a column reduction (deterministic order).
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.h"

namespace vmcx {

namespace {

constexpr int LP_T = 256;
constexpr int LP_CMAX = 32, LP_NMAX = 64, LP_EMAX = 2 * LP_NMAX;

struct LpOff {
  int W1, b1, W2, b2, W3, b3, A, bA, total;
};

__host__ __device__ inline LpOff lp_offsets(int C) {
  LpOff o;
  o.W1 = 0;
  o.b1 = o.W1 + C * 2 * 9;
  o.W2 = o.b1 + C;
  o.b2 = o.W2 + C * C * 9;
  o.W3 = o.b2 + C;
  o.b3 = o.W3 + C * C * 9;
  o.A = o.b3 + C;
  o.bA = o.A + 3 * 2 * C;
  o.total = o.bA + 3;
  return o;
}

// edge e of the sorted OBC list: (s, s+1) then (s, s+Ly) for s ascending
__host__ __device__ inline int lp_n_edges(int Lx, int Ly) {
  return Lx * (Ly - 1) + (Lx - 1) * Ly;
}

struct LpSh {
  float h[4][LP_CMAX * LP_NMAX];   // h0 (2 ch), h1, h2, h3 (post-activation)
  float z[3][LP_CMAX * LP_NMAX];   // pre-activations of layers 1..3
  int ei[LP_EMAX], ej[LP_EMAX];
  int x[LP_NMAX];
};

__device__ void lp_edges(int Lx, int Ly, int *ei, int *ej) {
  if (threadIdx.x == 0) {
    int e = 0;
    for (int s = 0; s < Lx * Ly; ++s) {
      const int xx = s / Ly, yy = s - xx * Ly;
      if (yy + 1 < Ly) { ei[e] = s; ej[e] = s + 1; ++e; }
      if (xx + 1 < Lx) { ei[e] = s; ej[e] = s + Ly; ++e; }
    }
  }
}

// out[co][s] = b[co] + sum_ci sum_taps W[co][ci][dx][dy] in[ci][s + (dx-1, dy-1)]
__device__ void lp_conv(int Lx, int Ly, int Cin, int Cout, const float *W, const float *b,
                        const float *in, float *z, float *h) {
  const int N = Lx * Ly;
  for (int idx = threadIdx.x; idx < Cout * N; idx += LP_T) {
    const int co = idx / N, s = idx - co * N;
    const int x = s / Ly, y = s - x * Ly;
    float acc = b[co];
    for (int ci = 0; ci < Cin; ++ci) {
      const float *w = W + (co * Cin + ci) * 9;
      const float *hi = in + ci * N;
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        const int xx = x + dx - 1;
        if (xx < 0 || xx >= Lx) continue;
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
          const int yy = y + dy - 1;
          if (yy < 0 || yy >= Ly) continue;
          acc += w[dx * 3 + dy] * hi[xx * Ly + yy];
        }
      }
    }
    z[idx] = acc;
    h[idx] = acc > 0.0f ? acc : 0.0f;
  }
  __syncthreads();
}

__device__ void lp_net(const float *P, int C, int Lx, int Ly, const int64_t *cfg, LpSh &S) {
  const int N = Lx * Ly;
  const LpOff o = lp_offsets(C);
  for (int s = threadIdx.x; s < N; s += LP_T) {
    const int v = (int)cfg[s];
    S.x[s] = v;
    S.h[0][s] = (float)((v >> 1) & 1);
    S.h[0][N + s] = (float)(v & 1);
  }
  lp_edges(Lx, Ly, S.ei, S.ej);
  __syncthreads();
  lp_conv(Lx, Ly, 2, C, P + o.W1, P + o.b1, S.h[0], S.z[0], S.h[1]);
  lp_conv(Lx, Ly, C, C, P + o.W2, P + o.b2, S.h[1], S.z[1], S.h[2]);
  lp_conv(Lx, Ly, C, C, P + o.W3, P + o.b3, S.h[2], S.z[2], S.h[3]);
}

__device__ __forceinline__ bool lp_valid(int t, int xi, int xj) {
  if (t == 0) return ((xi >> 1) & 1) != ((xj >> 1) & 1);
  if (t == 1) return (xi & 1) != (xj & 1);
  return (xi == 1 && xj == 2) || (xi == 2 && xj == 1);
}

}  // namespace

// logits[b][m] (FP64; -inf for invalid moves)
__global__ void __launch_bounds__(LP_T)
    lp_logits_kernel(const float *__restrict__ P, int C, int Lx, int Ly, int64_t B,
                     const int64_t *__restrict__ cfg, double *logits) {
  const int64_t b = blockIdx.x;
  if (b >= B) return;
  __shared__ LpSh S;
  const int N = Lx * Ly, E = lp_n_edges(Lx, Ly);
  lp_net(P, C, Lx, Ly, cfg + b * N, S);
  const LpOff o = lp_offsets(C);
  const float *A = P + o.A, *bA = P + o.bA;
  for (int m = threadIdx.x; m < 3 * E; m += LP_T) {
    const int e = m / 3, t = m - 3 * e;
    const int i = S.ei[e], j = S.ej[e];
    double v = -INFINITY;
    if (lp_valid(t, S.x[i], S.x[j])) {
      double acc = bA[t];
      for (int c = 0; c < C; ++c)
        acc += (double)A[t * 2 * C + c] * S.h[3][c * N + i] +
               (double)A[t * 2 * C + C + c] * S.h[3][c * N + j];
      v = acc;
    }
    logits[b * 3 * E + m] = v;
  }
}

// one thread per configuration: ln q of the chosen move (sample = 1: draw
// it by inverse CDF from u_choice / Philox and build the proposed config)
__global__ void lp_choose_kernel(int64_t B, int N, int Lx, int Ly, int E, int sample,
                                 const double *__restrict__ logits, uint64_t seed,
                                 int64_t step, const double *__restrict__ u_choice,
                                 const int64_t *__restrict__ cfg, int32_t *move,
                                 double *logq, int64_t *pcfg) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double *l = logits + b * 3 * E;
  double mx = -INFINITY;
  for (int m = 0; m < 3 * E; ++m) mx = fmax(mx, l[m]);
  if (mx == -INFINITY) {   // no valid move
    if (sample) {
      move[b] = -1;
      for (int s = 0; s < N; ++s) pcfg[b * N + s] = cfg[b * N + s];
    }
    logq[b] = 0.0;
    return;
  }
  double z = 0.0;
  for (int m = 0; m < 3 * E; ++m)
    if (l[m] != -INFINITY) z += exp(l[m] - mx);
  const double lse = mx + log(z);
  int m_sel;
  if (sample) {
    double u;
    if (u_choice) {
      u = u_choice[b];
    } else {
      const u32x4 r = philox4x32_10((uint32_t)step, (uint32_t)((uint64_t)step >> 32),
                                    (uint32_t)b, 0x4C50u, (uint32_t)seed,
                                    (uint32_t)(seed >> 32));
      u = u53(r.x, r.y);
    }
    double c = 0.0;
    m_sel = -1;
    for (int m = 0; m < 3 * E; ++m) {
      if (l[m] == -INFINITY) continue;
      c += exp(l[m] - lse);
      m_sel = m;
      if (u < c) break;
    }
    move[b] = m_sel;
    // proposed configuration
    int i = -1, j = -1;
    {
      const int e = m_sel / 3;
      int k = 0;
      for (int s = 0; s < N && i < 0; ++s) {
        const int xx = s / Ly, yy = s - xx * Ly;
        if (yy + 1 < Ly) { if (k == e) { i = s; j = s + 1; } ++k; }
        if (i < 0 && xx + 1 < Lx) { if (k == e) { i = s; j = s + Ly; } ++k; }
      }
    }
    const int t = m_sel - 3 * (m_sel / 3);
    for (int s = 0; s < N; ++s) pcfg[b * N + s] = cfg[b * N + s];
    const int64_t xi = cfg[b * N + i], xj = cfg[b * N + j];
    if (t == 0) {
      pcfg[b * N + i] = xi ^ 2;
      pcfg[b * N + j] = xj ^ 2;
    } else if (t == 1) {
      pcfg[b * N + i] = xi ^ 1;
      pcfg[b * N + j] = xj ^ 1;
    } else {
      pcfg[b * N + i] = xj;
      pcfg[b * N + j] = xi;
    }
  } else {
    m_sel = move[b];
  }
  logq[b] = (m_sel >= 0 && l[m_sel] != -INFINITY) ? l[m_sel] - lse : -INFINITY;
}

// Metropolis-Hastings test and commit
__global__ void lp_accept_kernel(int64_t B, int N, uint64_t seed, int64_t step,
                                 const double *__restrict__ u_acc,
                                 const int32_t *__restrict__ move,
                                 const int64_t *__restrict__ pcfg,
                                 const double *__restrict__ psign,
                                 const double *__restrict__ plog,
                                 const double *__restrict__ logq_f,
                                 const double *__restrict__ logq_r, int64_t *cfg,
                                 double *sign, double *logabs, int64_t *n_accept,
                                 uint8_t *accepted) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  if (accepted) accepted[b] = 0;
  if (move[b] < 0) return;
  double u;
  if (u_acc) {
    u = u_acc[b];
  } else {
    const u32x4 r = philox4x32_10((uint32_t)step, (uint32_t)((uint64_t)step >> 32), (uint32_t)b,
                                  0x4C50u, (uint32_t)seed, (uint32_t)(seed >> 32));
    u = u53(r.z, r.w);
  }
  const double ratio = exp(2.0 * (plog[b] - logabs[b]) + logq_r[b] - logq_f[b]);
  if (!(u < ratio)) return;
  for (int s = 0; s < N; ++s) cfg[b * N + s] = pcfg[b * N + s];
  sign[b] = psign[b];
  logabs[b] = plog[b];
  if (n_accept) n_accept[b] += 1;
  if (accepted) accepted[b] = 1;
}

// per-configuration gradient rows of sum_m dlogits[m] logits[m]
__global__ void __launch_bounds__(LP_T)
    lp_backward_kernel(const float *__restrict__ P, int C, int Lx, int Ly, int64_t B,
                       const int64_t *__restrict__ cfg, const double *__restrict__ dlogits,
                       float *G) {
  const int64_t b = blockIdx.x;
  if (b >= B) return;
  __shared__ LpSh S;
  __shared__ float dh[2][LP_CMAX * LP_NMAX];
  const int N = Lx * Ly, E = lp_n_edges(Lx, Ly);
  lp_net(P, C, Lx, Ly, cfg + b * N, S);
  const LpOff o = lp_offsets(C);
  const float *A = P + o.A;
  float *g = G + b * (int64_t)o.total;
  const double *dl = dlogits + b * 3 * E;
  // head: dA, dbA, dh3
  for (int idx = threadIdx.x; idx < 3 * 2 * C; idx += LP_T) {
    const int t = idx / (2 * C), k = idx - t * 2 * C;
    const int c = k < C ? k : k - C;
    double acc = 0.0;
    for (int e = 0; e < E; ++e) {
      const double d = dl[3 * e + t];
      if (d == 0.0) continue;
      acc += d * S.h[3][c * N + (k < C ? S.ei[e] : S.ej[e])];
    }
    g[o.A + idx] = (float)acc;
  }
  for (int t = threadIdx.x; t < 3; t += LP_T) {
    double acc = 0.0;
    for (int e = 0; e < E; ++e) acc += dl[3 * e + t];
    g[o.bA + t] = (float)acc;
  }
  for (int idx = threadIdx.x; idx < C * N; idx += LP_T) {
    const int c = idx / N, s = idx - c * N;
    double acc = 0.0;
    for (int e = 0; e < E; ++e) {
      const bool ii = S.ei[e] == s, jj = S.ej[e] == s;
      if (!ii && !jj) continue;
      for (int t = 0; t < 3; ++t) {
        const double d = dl[3 * e + t];
        if (ii) acc += d * A[t * 2 * C + c];
        if (jj) acc += d * A[t * 2 * C + C + c];
      }
    }
    // through the ReLU of layer 3
    dh[0][idx] = S.z[2][idx] > 0.0f ? (float)acc : 0.0f;
  }
  __syncthreads();
  // layers 3, 2, 1: dz (dh[cur]) -> dW, db, and dz of the layer below
  int cur = 0;
  for (int layer = 3; layer >= 1; --layer) {
    const int Cin = layer == 1 ? 2 : C;
    const int offW = layer == 3 ? o.W3 : (layer == 2 ? o.W2 : o.W1);
    const int offb = layer == 3 ? o.b3 : (layer == 2 ? o.b2 : o.b1);
    const float *hin = S.h[layer - 1];
    const float *dz = dh[cur];
    for (int idx = threadIdx.x; idx < C * Cin * 9; idx += LP_T) {
      const int tap = idx % 9, ci = (idx / 9) % Cin, co = idx / (9 * Cin);
      const int dx = tap / 3 - 1, dy = tap % 3 - 1;
      float acc = 0.0f;
      for (int s = 0; s < N; ++s) {
        const int x = s / Ly + dx, y = s % Ly + dy;
        if (x < 0 || x >= Lx || y < 0 || y >= Ly) continue;
        acc += dz[co * N + s] * hin[ci * N + x * Ly + y];
      }
      g[offW + idx] = acc;
    }
    for (int co = threadIdx.x; co < C; co += LP_T) {
      float acc = 0.0f;
      for (int s = 0; s < N; ++s) acc += dz[co * N + s];
      g[offb + co] = acc;
    }
    if (layer > 1) {
      const float *W = P + offW;
      float *dn = dh[cur ^ 1];
      for (int idx = threadIdx.x; idx < Cin * N; idx += LP_T) {
        const int ci = idx / N, s = idx - ci * N;
        const int x = s / Ly, y = s - x * Ly;
        float acc = 0.0f;
        for (int co = 0; co < C; ++co)
          for (int dx = 0; dx < 3; ++dx) {
            const int xo = x - (dx - 1);
            if (xo < 0 || xo >= Lx) continue;
            for (int dy = 0; dy < 3; ++dy) {
              const int yo = y - (dy - 1);
              if (yo < 0 || yo >= Ly) continue;
              acc += W[(co * Cin + ci) * 9 + dx * 3 + dy] * dz[co * N + xo * Ly + yo];
            }
          }
        dn[idx] = S.z[layer - 2][idx] > 0.0f ? acc : 0.0f;
      }
      __syncthreads();
      cur ^= 1;
    }
  }
}

// out[k] = sum_b G[b][k] (FP64 accumulation, fixed order)
__global__ void lp_colsum_kernel(int64_t B, int64_t n, const float *__restrict__ G,
                                 float *out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double acc = 0.0;
  for (int64_t b = 0; b < B; ++b) acc += G[b * n + k];
  out[k] = (float)acc;
}

// every move applied to each configuration: out[b][m][N] (a copy of x for
// invalid moves)
__global__ void lp_candidates_kernel(int64_t B, int Lx, int Ly, const int64_t *__restrict__ cfg,
                                     int64_t *out) {
  const int N = Lx * Ly, E = lp_n_edges(Lx, Ly), M = 3 * E;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * M) return;
  const int64_t b = idx / M;
  const int m = (int)(idx - b * M), e = m / 3, t = m - 3 * e;
  int i = -1, j = -1, k = 0;
  for (int s = 0; s < N && i < 0; ++s) {
    const int xx = s / Ly, yy = s - xx * Ly;
    if (yy + 1 < Ly) { if (k == e) { i = s; j = s + 1; } ++k; }
    if (i < 0 && xx + 1 < Lx) { if (k == e) { i = s; j = s + Ly; } ++k; }
  }
  const int64_t *c = cfg + b * N;
  int64_t *o = out + idx * N;
  for (int s = 0; s < N; ++s) o[s] = c[s];
  const int xi = (int)c[i], xj = (int)c[j];
  if (!lp_valid(t, xi, xj)) return;
  if (t == 0) { o[i] = xi ^ 2; o[j] = xj ^ 2; }
  else if (t == 1) { o[i] = xi ^ 1; o[j] = xj ^ 1; }
  else { o[i] = xj; o[j] = xi; }
}

__global__ void lp_axpy_kernel(int64_t n, float a, const float *__restrict__ x, float *y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

// ------------------------------------------------------------ host side
static int lp_check(int C, int Lx, int Ly) {
  VMCX_CHECK_ARG(C >= 1 && C <= LP_CMAX, "C must be 1..32");
  VMCX_CHECK_ARG(Lx >= 1 && Ly >= 1 && Lx * Ly <= LP_NMAX, "lattice must have <= 64 sites");
  VMCX_CHECK_ARG(Lx * Ly >= 2, "lattice needs an edge");
  return 0;
}

int lp_n_params(int C) { return (C >= 1 && C <= LP_CMAX) ? lp_offsets(C).total : -1; }

int lp_n_moves(int Lx, int Ly) {
  if (Lx < 1 || Ly < 1 || Lx * Ly > LP_NMAX) return -1;
  return 3 * lp_n_edges(Lx, Ly);
}

int lp_logits(const float *P, int C, int Lx, int Ly, int64_t B, const int64_t *cfg,
              double *logits, hipStream_t st) {
  if (lp_check(C, Lx, Ly)) return 1;
  VMCX_CHECK_ARG(B >= 0, "B < 0");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(P && cfg && logits, "null pointer");
  hipLaunchKernelGGL(lp_logits_kernel, dim3((unsigned)B), dim3(LP_T), 0, st, P, C, Lx, Ly, B,
                     cfg, logits);
  VMCX_LAUNCHED();
  return 0;
}

int lp_propose(const float *P, int C, int Lx, int Ly, int64_t B, const int64_t *cfg,
               uint64_t seed, int64_t step, const double *u_choice, double *logits,
               int32_t *move, double *logq_f, int64_t *pcfg, double *logq_r,
               hipStream_t st) {
  int rc = lp_logits(P, C, Lx, Ly, B, cfg, logits, st);
  if (rc || B == 0) return rc;
  VMCX_CHECK_ARG(move && logq_f && pcfg && logq_r, "null pointer");
  const int N = Lx * Ly, E = lp_n_edges(Lx, Ly);
  const unsigned g = (unsigned)((B + 255) / 256);
  hipLaunchKernelGGL(lp_choose_kernel, dim3(g), dim3(256), 0, st, B, N, Lx, Ly, E, 1, logits,
                     seed, step, u_choice, cfg, move, logq_f, pcfg);
  VMCX_LAUNCHED();
  // reverse probability: the same move evaluated at the proposed configuration
  rc = lp_logits(P, C, Lx, Ly, B, pcfg, logits, st);
  if (rc) return rc;
  hipLaunchKernelGGL(lp_choose_kernel, dim3(g), dim3(256), 0, st, B, N, Lx, Ly, E, 0, logits,
                     seed, step, (const double *)nullptr, pcfg, move, logq_r,
                     (int64_t *)nullptr);
  VMCX_LAUNCHED();
  return 0;
}

int lp_accept(int64_t B, int N, uint64_t seed, int64_t step, const double *u_acc,
              const int32_t *move, const int64_t *pcfg, const double *psign,
              const double *plog, const double *logq_f, const double *logq_r, int64_t *cfg,
              double *sign, double *logabs, int64_t *n_accept, uint8_t *accepted,
              hipStream_t st) {
  VMCX_CHECK_ARG(B >= 0 && N >= 1 && N <= LP_NMAX, "bad sizes");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(move && pcfg && psign && plog && logq_f && logq_r && cfg && sign && logabs,
                 "null pointer");
  hipLaunchKernelGGL(lp_accept_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, st, B,
                     N, seed, step, u_acc, move, pcfg, psign, plog, logq_f, logq_r, cfg, sign,
                     logabs, n_accept, accepted);
  VMCX_LAUNCHED();
  return 0;
}

int64_t lp_backward_workspace_bytes(int C, int64_t B) {
  if (C < 1 || C > LP_CMAX || B < 0) return -1;
  return (int64_t)B * lp_offsets(C).total * (int64_t)sizeof(float);
}

int lp_backward(const float *P, int C, int Lx, int Ly, int64_t B, const int64_t *cfg,
                const double *dlogits, void *ws, int64_t ws_bytes, float *grad,
                hipStream_t st) {
  if (lp_check(C, Lx, Ly)) return 1;
  VMCX_CHECK_ARG(B >= 0, "B < 0");
  const int np = lp_offsets(C).total;
  VMCX_CHECK_ARG(ws_bytes >= (int64_t)B * np * (int64_t)sizeof(float), "workspace too small");
  VMCX_CHECK_ARG(P && grad && (B == 0 || (cfg && dlogits && ws)), "null pointer");
  if (B > 0) {
    hipLaunchKernelGGL(lp_backward_kernel, dim3((unsigned)B), dim3(LP_T), 0, st, P, C, Lx, Ly,
                       B, cfg, dlogits, (float *)ws);
    VMCX_LAUNCHED();
  }
  hipLaunchKernelGGL(lp_colsum_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, B,
                     (int64_t)np, (const float *)ws, grad);
  VMCX_LAUNCHED();
  return 0;
}

int lp_candidates(int64_t B, int Lx, int Ly, const int64_t *cfg, int64_t *out,
                  hipStream_t st) {
  VMCX_CHECK_ARG(B >= 0 && Lx >= 1 && Ly >= 1 && Lx * Ly <= LP_NMAX && Lx * Ly >= 2,
                 "bad sizes");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(cfg && out, "null pointer");
  const int64_t n = B * 3 * lp_n_edges(Lx, Ly);
  hipLaunchKernelGGL(lp_candidates_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     B, Lx, Ly, cfg, out);
  VMCX_LAUNCHED();
  return 0;
}

int lp_axpy(int64_t n, float a, const float *x, float *y, hipStream_t st) {
  VMCX_CHECK_ARG(n >= 0 && (n == 0 || (x && y)), "bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(lp_axpy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a,
                     x, y);
  VMCX_LAUNCHED();
  return 0;
}

}  // namespace vmcx
