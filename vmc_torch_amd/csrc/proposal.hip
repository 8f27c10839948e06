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

// the sorted edge list built by all threads of the block: site s = x Ly + y owns
// the edges (s, s + 1) (y + 1 < Ly) then (s, s + Ly) (x + 1 < Lx), which
// start at x (2 Ly - 1) + y (1 + [x + 1 < Lx]) in the sorted order
__device__ void lp_edges_par(int Lx, int Ly, int *ei, int *ej) {
  for (int s = threadIdx.x; s < Lx * Ly; s += blockDim.x) {
    const int x = s / Ly, y = s - x * Ly;
    const bool down = x + 1 < Lx;
    int e = x * (2 * Ly - 1) + y * (down ? 2 : 1);
    if (y + 1 < Ly) { ei[e] = s; ej[e] = s + 1; ++e; }
    if (down) { ei[e] = s; ej[e] = s + Ly; }
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
  lp_edges_par(Lx, Ly, S.ei, S.ej);
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

// Forward pass batched over configurations: G configurations per workgroup,
// one thread per (configuration, site).  Each thread accumulates CB output
// channels of its site at a time; the weights are workgroup-uniform (scalar
// loads, one FMA per weight and thread), the activations ping-pong between
// two LDS buffers laid out [channel][G * (N + 1)] with a zero slot per
// configuration that the padding taps read.  The sums run in the order of
// the per-configuration form (bias, then input channel, then tap), so the
// logits are the same.
constexpr int LPB_SLOTS = 240;   // G * (N + 1) <= LPB_SLOTS (LDS <= 64 KiB), G * N <= threads

// configurations per workgroup for a block of `threads` threads
__host__ __device__ inline int lpb_group(int N, int threads) {
  int g = threads / N;
  if (g * (N + 1) > LPB_SLOTS) g = LPB_SLOTS / (N + 1);
  return g < 1 ? 1 : g;
}

__host__ __device__ inline size_t lpb_lds_bytes(int C, int N, int G) {
  const int R = G * (N + 1);
  return (size_t)(2 * C * R + R + 2 * LP_EMAX) * sizeof(float);
}

template <int CB>
__global__ void __launch_bounds__(LP_T)
    lp_logits_batched_kernel(const float *__restrict__ P, int C, int Lx, int Ly, int G,
                             int64_t B, const int64_t *__restrict__ cfg,
                             double *__restrict__ logits) {
  extern __shared__ float lpb_sh[];
  const int N = Lx * Ly, E = lp_n_edges(Lx, Ly), R = G * (N + 1);
  float *hb0 = lpb_sh, *hb1 = lpb_sh + C * R;
  int *xs = reinterpret_cast<int *>(lpb_sh + 2 * C * R);
  int *ei = xs + R, *ej = ei + LP_EMAX;
  const int64_t b0 = (int64_t)blockIdx.x * G;
  const int t = threadIdx.x;
  const bool site = t < G * N;             // this thread owns a (configuration, site)
  const int g = site ? t / N : 0, s = site ? t - (t / N) * N : 0;
  // input channels (n_up, n_dn); configurations past B read as empty
  for (int idx = t; idx < R; idx += blockDim.x) {
    const int gg = idx / (N + 1), ss = idx - gg * (N + 1);
    int v = 0;
    if (ss < N && b0 + gg < B) v = (int)cfg[(b0 + gg) * N + ss];
    xs[idx] = v;
    hb0[idx] = (float)((v >> 1) & 1);
    hb0[R + idx] = (float)(v & 1);
  }
  // zero slots of every channel in both buffers
  for (int idx = t; idx < C * G; idx += blockDim.x) {
    const int c = idx / G, gg = idx - c * G;
    hb0[c * R + gg * (N + 1) + N] = 0.0f;
    hb1[c * R + gg * (N + 1) + N] = 0.0f;
  }
  lp_edges_par(Lx, Ly, ei, ej);
  int off[9];
  {
    const int x = s / Ly, y = s - (s / Ly) * Ly, base = g * (N + 1);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      const int xx = x + q / 3 - 1, yy = y + q % 3 - 1;
      off[q] = base + ((xx >= 0 && xx < Lx && yy >= 0 && yy < Ly) ? xx * Ly + yy : N);
    }
  }
  __syncthreads();
  const LpOff o = lp_offsets(C);
  const float *hin = hb0;
  float *hout = hb1;
  for (int layer = 0; layer < 3; ++layer) {
    const int Cin = layer == 0 ? 2 : C;
    const float *W = P + (layer == 0 ? o.W1 : (layer == 1 ? o.W2 : o.W3));
    const float *bias = P + (layer == 0 ? o.b1 : (layer == 1 ? o.b2 : o.b3));
    if (site) {
      for (int co0 = 0; co0 < C; co0 += CB) {
        float acc[CB];
#pragma unroll
        for (int k = 0; k < CB; ++k) acc[k] = co0 + k < C ? bias[co0 + k] : 0.0f;
        for (int ci = 0; ci < Cin; ++ci) {
          float v[9];
#pragma unroll
          for (int q = 0; q < 9; ++q) v[q] = hin[ci * R + off[q]];
#pragma unroll
          for (int k = 0; k < CB; ++k) {
            if (co0 + k < C) {
              const float *w = W + ((co0 + k) * Cin + ci) * 9;
#pragma unroll
              for (int q = 0; q < 9; ++q) acc[k] += w[q] * v[q];
            }
          }
        }
#pragma unroll
        for (int k = 0; k < CB; ++k)
          if (co0 + k < C) hout[(co0 + k) * R + off[4]] = acc[k] > 0.0f ? acc[k] : 0.0f;
      }
    }
    __syncthreads();
    const float *tmp = hout;
    hout = const_cast<float *>(hin);
    hin = tmp;
  }
  // head: one thread per (configuration, move), FP64 accumulation
  const float *A = P + o.A, *bA = P + o.bA;
  for (int idx = t; idx < G * 3 * E; idx += blockDim.x) {
    const int gg = idx / (3 * E), m = idx - gg * 3 * E;
    if (b0 + gg >= B) break;
    const int e = m / 3, tt = m - 3 * e, base = gg * (N + 1);
    const int i = base + ei[e], j = base + ej[e];
    double v = -INFINITY;
    if (lp_valid(tt, xs[i], xs[j])) {
      double acc = bA[tt];
      for (int c = 0; c < C; ++c)
        acc += (double)A[tt * 2 * C + c] * hin[c * R + i] +
               (double)A[tt * 2 * C + C + c] * hin[c * R + j];
      v = acc;
    }
    logits[(b0 + gg) * 3 * E + m] = v;
  }
}

// BF16 matrix-core forward (C = 32).  Layer 1 (2 -> 32 channels) runs in
// FP32 on the vector ALUs; layers 2 and 3 are implicit GEMMs
//   z[co][col] = sum_{tap, ci} W[co][ci][tap] h[ci][nbr(col, tap)]
// over the columns col = (configuration, site) of the workgroup, 32 x 32
// output tiles of v_mfma_f32_32x32x16_bf16 (FP32 accumulation), one K chunk
// = one tap and 16 input channels.  Activations are stored BF16 and
// channel-minor ([slot][32], one zero slot per configuration) so that the 8
// channels a lane contributes to a chunk are one 16-byte LDS read; the
// layer's weights are staged BF16 in LDS in fragment order (chunk, lane, 8).
// Bias and ReLU in FP32, the result rounded to BF16; the head accumulates in
// FP64 from the BF16 activations, as two per-site projections per move type
// (A[t, :C] . h3[i] and A[t, C:] . h3[j], the oracle's grouping).
typedef __bf16 lp_bf8 __attribute__((ext_vector_type(8)));
typedef float lp_f16v __attribute__((ext_vector_type(16)));

constexpr int LPM_T = 256;         // 4 waves
constexpr int LPM_COLS = 256;      // at most 8 column tiles per workgroup
constexpr int LPM_RMAX = 288;      // slots G * (N + 1)
constexpr int LPM_CH = 32;
constexpr int LPM_KCH = 18;        // 9 taps x 2 halves of the 32 input channels

__host__ __device__ inline int lpm_group(int N) {
  int g = LPM_COLS / N;
  if (g * (N + 1) > LPM_RMAX) g = LPM_RMAX / (N + 1);
  return g < 1 ? 1 : g;
}

__global__ void __launch_bounds__(LPM_T)
    lp_logits_mfma_kernel(const float *__restrict__ P, int Lx, int Ly, int G, int64_t B,
                          const int64_t *__restrict__ cfg, double *__restrict__ logits) {
  __shared__ __attribute__((aligned(16))) __bf16 act[2][LPM_RMAX * LPM_CH];
  __shared__ __attribute__((aligned(16))) __bf16 wf[2][LPM_KCH * 64 * 8];
  __shared__ int xs[LPM_RMAX];
  __shared__ int ei[LP_EMAX], ej[LP_EMAX];
  __shared__ float w1s[LPM_CH * 2 * 9 + LPM_CH];       // layer 1 weights, then biases
  static_assert(sizeof(double) * LPM_RMAX * 6 <= sizeof(__bf16) * LPM_RMAX * LPM_CH,
                "the head projections alias a layer buffer");
  // the head projections per slot, in act[1] once layer 3 has read it
  double *proj = reinterpret_cast<double *>(act[1]);
  const int N = Lx * Ly, E = lp_n_edges(Lx, Ly), R = G * (N + 1), ncol = G * N;
  const int64_t b0 = (int64_t)blockIdx.x * G;
  const int t = threadIdx.x;
  const LpOff o = lp_offsets(LPM_CH);
  for (int idx = t; idx < LPM_CH * 2 * 9 + LPM_CH; idx += LPM_T)
    w1s[idx] = idx < LPM_CH * 2 * 9 ? P[o.W1 + idx] : P[o.b1 + idx - LPM_CH * 2 * 9];
  // layers 2 and 3 weights as BF16 matrix-core fragments, one 16-byte
  // fragment per step: chunk = tap * 2 + (ci >> 4), lane ln = co + 32
  // ((ci & 15) >> 3), the 8 input channels ci & 7
  for (int f = t; f < 2 * LPM_KCH * 64; f += LPM_T) {
    const int l = f / (LPM_KCH * 64), ff = f - l * (LPM_KCH * 64);
    const float *W = P + (l == 0 ? o.W2 : o.W3);
    const int chunk = ff >> 6, ln = ff & 63;
    const int co = ln & 31, tap = chunk >> 1;
    const int ci0 = 16 * (chunk & 1) + 8 * (ln >> 5);
    lp_bf8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (__bf16)W[(co * LPM_CH + ci0 + e) * 9 + tap];
    *reinterpret_cast<lp_bf8 *>(&wf[l][ff * 8]) = v;
  }
  for (int idx = t; idx < R; idx += LPM_T) {
    const int gg = idx / (N + 1), ss = idx - gg * (N + 1);
    xs[idx] = (ss < N && b0 + gg < B) ? (int)cfg[(b0 + gg) * N + ss] : 0;
  }
  for (int idx = t; idx < G * LPM_CH; idx += LPM_T) {
    const int gg = idx / LPM_CH, c = idx - gg * LPM_CH;
    act[0][(gg * (N + 1) + N) * LPM_CH + c] = (__bf16)0.0f;
    act[1][(gg * (N + 1) + N) * LPM_CH + c] = (__bf16)0.0f;
  }
  lp_edges_par(Lx, Ly, ei, ej);
  __syncthreads();
  // layer 1, FP32: thread = (column, half of the output channels), weights
  // from LDS
  {
    const float *W1 = w1s, *b1 = w1s + LPM_CH * 2 * 9;
    for (int idx = t; idx < 2 * ncol; idx += LPM_T) {
      const int col = idx >> 1, half = idx & 1;
      const int g = col / N, s = col - g * N, x = s / Ly, y = s - (s / Ly) * Ly;
      const int base = g * (N + 1);
      float hu[9], hd[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        const int xx = x + q / 3 - 1, yy = y + q % 3 - 1;
        const int v = xs[base + ((xx >= 0 && xx < Lx && yy >= 0 && yy < Ly) ? xx * Ly + yy : N)];
        hu[q] = (float)((v >> 1) & 1);
        hd[q] = (float)(v & 1);
      }
      __bf16 *out = act[0] + (base + s) * LPM_CH;
      for (int k = 0; k < 16; ++k) {
        const int co = half * 16 + k;
        float acc = b1[co];
#pragma unroll
        for (int q = 0; q < 9; ++q) acc += W1[(co * 2) * 9 + q] * hu[q];
#pragma unroll
        for (int q = 0; q < 9; ++q) acc += W1[(co * 2 + 1) * 9 + q] * hd[q];
        out[co] = (__bf16)(acc > 0.0f ? acc : 0.0f);
      }
    }
  }
  const int wave = t >> 6, lane = t & 63, kh = lane >> 5;
  for (int layer = 1; layer < 3; ++layer) {
    const float *bias = P + (layer == 1 ? o.b2 : o.b3);
    const __bf16 *in = act[layer == 1 ? 0 : 1];
    __bf16 *outb = act[layer == 1 ? 1 : 0];
    const __bf16 *wl = wf[layer - 1];
    __syncthreads();   // the previous layer's writes of `in` are done
    for (int tile = wave; tile * 32 < ncol; tile += LPM_T / 64) {
      const int col = tile * 32 + (lane & 31);
      const bool ok = col < ncol;
      const int g = ok ? col / N : 0, s = ok ? col - g * N : 0;
      const int x = s / Ly, y = s - (s / Ly) * Ly, base = g * (N + 1);
      int nb[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        const int xx = x + q / 3 - 1, yy = y + q % 3 - 1;
        nb[q] = base + ((ok && xx >= 0 && xx < Lx && yy >= 0 && yy < Ly) ? xx * Ly + yy : N);
      }
      lp_f16v acc;
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
#pragma unroll
      for (int q = 0; q < 9; ++q)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const lp_bf8 a = *reinterpret_cast<const lp_bf8 *>(&wl[((q * 2 + h) * 64 + lane) * 8]);
          const lp_bf8 bv =
              *reinterpret_cast<const lp_bf8 *>(&in[nb[q] * LPM_CH + h * 16 + kh * 8]);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv, acc, 0, 0, 0);
        }
      if (ok) {
        // D lane l, register r: row (output channel) 8 (r / 4) + 4 (l / 32) + r % 4
        __bf16 *out = outb + (base + s) * LPM_CH;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int co = 8 * (r >> 2) + 4 * kh + (r & 3);
          const float v = acc[r] + bias[co];
          out[co] = (__bf16)(v > 0.0f ? v : 0.0f);
        }
      }
    }
  }
  __syncthreads();
  // head in FP64: per site and move type the two projections
  // proj[slot][2 tt + side] = A[tt, side C .. side C + C - 1] . h3[slot], then
  // logit(e, tt) = (bA[tt] + proj[i][2 tt]) + proj[j][2 tt + 1]
  const float *A = P + o.A, *bA = P + o.bA;
  const __bf16 *h3 = act[0];
  for (int idx = t; idx < ncol * 6; idx += LPM_T) {
    const int col = idx / 6, k = idx - col * 6;
    const int g = col / N, slot = col + g;    // g (N + 1) + s
    const float *a = A + (k >> 1) * 2 * LPM_CH + (k & 1) * LPM_CH;
    const __bf16 *hs = h3 + slot * LPM_CH;
    double acc = 0.0;
#pragma unroll 8
    for (int c = 0; c < LPM_CH; ++c) acc += (double)a[c] * (double)(float)hs[c];
    proj[slot * 6 + k] = acc;
  }
  __syncthreads();
  for (int idx = t; idx < G * 3 * E; idx += LPM_T) {
    const int gg = idx / (3 * E), m = idx - gg * 3 * E;
    if (b0 + gg >= B) break;
    const int e = m / 3, tt = m - 3 * e, base = gg * (N + 1);
    const int i = base + ei[e], j = base + ej[e];
    double v = -INFINITY;
    if (lp_valid(tt, xs[i], xs[j]))
      v = ((double)bA[tt] + proj[i * 6 + 2 * tt]) + proj[j * 6 + 2 * tt + 1];
    logits[(b0 + gg) * 3 * E + m] = v;
  }
}

// one wave per configuration: ln q of the chosen move (sample = 1: draw it
// by inverse CDF from u_choice / Philox and build the proposed
// configuration).  Lane l owns the moves [l * per, (l + 1) * per); the
// maximum and the normaliser are butterfly reductions (every lane ends with
// the same bits), the CDF is an inclusive scan of the per-lane masses and
// the lane whose running total first passes u scans its own moves in order.
constexpr int LPC_WAVES = 4;

__global__ void __launch_bounds__(64 * LPC_WAVES)
    lp_choose_kernel(int64_t B, int N, int Lx, int Ly, int E, int sample,
                     const double *__restrict__ logits, uint64_t seed, int64_t step0,
                     const double *__restrict__ u_choice, const int64_t *__restrict__ cfg,
                     int32_t *move, double *logq, int64_t *pcfg,
                     const int64_t *__restrict__ step_dev) {
  const int64_t step = step0 + (step_dev ? step_dev[0] : 0);
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * LPC_WAVES + (threadIdx.x >> 6);
  if (b >= B) return;   // whole wave
  const int M = 3 * E, per = (M + 63) / 64;
  const int m0 = min(M, lane * per), m1 = min(M, m0 + per);
  const double *l = logits + b * M;
  double mx = -INFINITY;
  for (int m = m0; m < m1; ++m) mx = fmax(mx, l[m]);
  for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  const int64_t *c = cfg + b * N;
  if (mx == -INFINITY) {   // no valid move
    if (sample) {
      if (lane == 0) move[b] = -1;
      if (lane < N) pcfg[b * N + lane] = c[lane];
    }
    if (lane == 0) logq[b] = 0.0;
    return;
  }
  double z = 0.0;
  for (int m = m0; m < m1; ++m)
    if (l[m] != -INFINITY) z += exp(l[m] - mx);
  for (int o = 1; o < 64; o <<= 1) z += __shfl_xor(z, o);
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
    double pl = 0.0;
    for (int m = m0; m < m1; ++m)
      if (l[m] != -INFINITY) pl += exp(l[m] - lse);
    double inc = pl;
    for (int o = 1; o < 64; o <<= 1) {
      const double v = __shfl_up(inc, o);
      if (lane >= o) inc += v;
    }
    double exc = __shfl_up(inc, 1);
    if (lane == 0) exc = 0.0;
    const unsigned long long pass = __ballot(u < inc);
    int sel = -1, mine = -1;
    if (pass) {
      sel = __ffsll((long long)pass) - 1;
      if (lane == sel) {
        double cc = exc;
        for (int m = m0; m < m1; ++m) {
          if (l[m] == -INFINITY) continue;
          cc += exp(l[m] - lse);
          mine = m;
          if (u < cc) break;
        }
      }
    } else {   // u beyond the rounded total: the last valid move
      const unsigned long long nz = __ballot(pl > 0.0);
      sel = 63 - __clzll((long long)nz);
      if (lane == sel)
        for (int m = m0; m < m1; ++m)
          if (l[m] != -INFINITY) mine = m;
    }
    m_sel = __shfl(mine, sel);
    if (lane == 0) move[b] = m_sel;
    // edge e = (i, j): i is the last site whose first edge index
    // k_s = x (2 Ly - 1) + y + y [x + 1 < Lx] is <= e
    const int e = m_sel / 3, t = m_sel - 3 * e;
    int ks = 0, v = 0;
    if (lane < N) {
      const int x = lane / Ly, y = lane - x * Ly;
      ks = x * (2 * Ly - 1) + y + (x + 1 < Lx ? y : 0);
      v = (int)c[lane];
    }
    const int i = 63 - __clzll((long long)__ballot(lane < N && ks <= e));
    const int ki = __shfl(ks, i);
    const int j = (e == ki && (i % Ly) + 1 < Ly) ? i + 1 : i + Ly;
    const int xi = __shfl(v, i), xj = __shfl(v, j);
    if (lane == i) v = t == 0 ? xi ^ 2 : (t == 1 ? xi ^ 1 : xj);
    if (lane == j) v = t == 0 ? xj ^ 2 : (t == 1 ? xj ^ 1 : xi);
    if (lane < N) pcfg[b * N + lane] = v;
  } else {
    m_sel = move[b];
  }
  if (lane == 0)
    logq[b] = (m_sel >= 0 && l[m_sel] != -INFINITY) ? l[m_sel] - lse : -INFINITY;
}

// Metropolis-Hastings test and commit
__global__ void lp_accept_kernel(int64_t B, int N, uint64_t seed, int64_t step0,
                                 const double *__restrict__ u_acc,
                                 const int32_t *__restrict__ move,
                                 const int64_t *__restrict__ pcfg,
                                 const double *__restrict__ psign,
                                 const double *__restrict__ plog,
                                 const double *__restrict__ logq_f,
                                 const double *__restrict__ logq_r, int64_t *cfg,
                                 double *sign, double *logabs, int64_t *n_accept,
                                 uint8_t *accepted, const int64_t *__restrict__ step_dev) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int64_t step = step0 + (step_dev ? step_dev[0] : 0);
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

// lp_accept_kernel, lp_commit_logits_kernel and a masked row copy in one
// launch: one wave per walker, lane 0 takes the Metropolis-Hastings decision
// (the same arithmetic and draw), then the wave copies the accepted walker's
// configuration, logits row (logits_p -> logits_x) and state row (z1p -> z1,
// row_bytes, a multiple of 4; z1 may be NULL)
__global__ void __launch_bounds__(256)
    lp_accept_commit_kernel(int64_t B, int N, uint64_t seed, int64_t step0,
                            const double *__restrict__ u_acc, const int32_t *__restrict__ move,
                            const int64_t *__restrict__ pcfg, const double *__restrict__ psign,
                            const double *__restrict__ plog, const double *__restrict__ logq_f,
                            const double *__restrict__ logq_r, int64_t *cfg, double *sign,
                            double *logabs, int64_t *n_accept, uint8_t *accepted,
                            const int64_t *__restrict__ step_dev, int M,
                            const double *__restrict__ logits_p, double *logits_x,
                            int64_t row_words, const uint32_t *__restrict__ z1p, uint32_t *z1) {
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  int acc = 0;
  if (lane == 0) {
    const int64_t step = step0 + (step_dev ? step_dev[0] : 0);
    if (move[b] >= 0) {
      double u;
      if (u_acc) {
        u = u_acc[b];
      } else {
        const u32x4 r = philox4x32_10((uint32_t)step, (uint32_t)((uint64_t)step >> 32),
                                      (uint32_t)b, 0x4C50u, (uint32_t)seed,
                                      (uint32_t)(seed >> 32));
        u = u53(r.z, r.w);
      }
      const double ratio = exp(2.0 * (plog[b] - logabs[b]) + logq_r[b] - logq_f[b]);
      acc = u < ratio;
    }
    if (acc) {
      sign[b] = psign[b];
      logabs[b] = plog[b];
      if (n_accept) n_accept[b] += 1;
    }
    if (accepted) accepted[b] = (uint8_t)acc;
  }
  acc = __shfl(acc, 0, 64);
  if (!acc) return;
  for (int s = lane; s < N; s += 64) cfg[b * N + s] = pcfg[b * N + s];
  for (int m = lane; m < M; m += 64) logits_x[b * M + m] = logits_p[b * M + m];
  if (z1)
    for (int64_t k = lane; k < row_words; k += 64) z1[b * row_words + k] = z1p[b * row_words + k];
}

int lp_accept_commit(int64_t B, int N, uint64_t seed, int64_t step, const double *u_acc,
                     const int32_t *move, const int64_t *pcfg, const double *psign,
                     const double *plog, const double *logq_f, const double *logq_r,
                     int64_t *cfg, double *sign, double *logabs, int64_t *n_accept,
                     uint8_t *accepted, const int64_t *step_dev, int M, const double *logits_p,
                     double *logits_x, int64_t row_bytes, const void *z1p, void *z1,
                     hipStream_t st) {
  VMCX_CHECK_ARG(B >= 0 && N >= 1 && N <= LP_NMAX && M >= 1, "bad sizes");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(move && pcfg && psign && plog && logq_f && logq_r && cfg && sign && logabs &&
                     logits_p && logits_x,
                 "null pointer");
  VMCX_CHECK_ARG(z1 == nullptr || (z1p != nullptr && row_bytes > 0 && row_bytes % 4 == 0),
                 "state rows need z1p and a row size that is a multiple of 4 bytes");
  hipLaunchKernelGGL(lp_accept_commit_kernel, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, st,
                     B, N, seed, step, u_acc, move, pcfg, psign, plog, logq_f, logq_r, cfg, sign,
                     logabs, n_accept, accepted, step_dev, M, logits_p, logits_x,
                     z1 ? row_bytes / 4 : 0, (const uint32_t *)z1p, (uint32_t *)z1);
  VMCX_LAUNCHED();
  return 0;
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
              double *logits, int precision, hipStream_t st) {
  if (lp_check(C, Lx, Ly)) return 1;
  VMCX_CHECK_ARG(B >= 0, "B < 0");
  VMCX_CHECK_ARG(precision == 0 || precision == 1, "precision must be 0 (fp32) or 1 (bf16)");
  VMCX_CHECK_ARG(precision == 0 || C == LPM_CH, "the bf16 net needs C = 32");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(P && cfg && logits, "null pointer");
  const int N = Lx * Ly;
  if (precision == 1) {
    const int G = lpm_group(N);
    hipLaunchKernelGGL(lp_logits_mfma_kernel, dim3((unsigned)((B + G - 1) / G)), dim3(LPM_T), 0,
                       st, P, Lx, Ly, G, B, cfg, logits);
    VMCX_LAUNCHED();
    return 0;
  }
  // smallest block (64, 128, 256 threads) that leaves at most 1/8 of its
  // lanes without a site
  int threads = LP_T;
  for (int T = 64; T < LP_T; T *= 2)
    if (8 * lpb_group(N, T) * N >= 7 * T) { threads = T; break; }
  const int G = lpb_group(N, threads);
  const unsigned grid = (unsigned)((B + G - 1) / G);
  const size_t lds = lpb_lds_bytes(C, N, G);
  if (C > 8)
    hipLaunchKernelGGL(lp_logits_batched_kernel<16>, dim3(grid), dim3(threads), lds, st, P, C,
                       Lx, Ly, G, B, cfg, logits);
  else
    hipLaunchKernelGGL(lp_logits_batched_kernel<8>, dim3(grid), dim3(threads), lds, st, P, C,
                       Lx, Ly, G, B, cfg, logits);
  VMCX_LAUNCHED();
  return 0;
}

int lp_propose(const float *P, int C, int Lx, int Ly, int64_t B, const int64_t *cfg,
               uint64_t seed, int64_t step, const double *u_choice, double *logits,
               int32_t *move, double *logq_f, int64_t *pcfg, double *logq_r, int precision,
               hipStream_t st) {
  int rc = lp_logits(P, C, Lx, Ly, B, cfg, logits, precision, st);
  if (rc || B == 0) return rc;
  VMCX_CHECK_ARG(move && logq_f && pcfg && logq_r, "null pointer");
  const int N = Lx * Ly, E = lp_n_edges(Lx, Ly);
  const unsigned g = (unsigned)((B + LPC_WAVES - 1) / LPC_WAVES);
  hipLaunchKernelGGL(lp_choose_kernel, dim3(g), dim3(64 * LPC_WAVES), 0, st, B, N, Lx, Ly, E, 1,
                     logits, seed, step, u_choice, cfg, move, logq_f, pcfg,
                     (const int64_t *)nullptr);
  VMCX_LAUNCHED();
  // reverse probability: the same move evaluated at the proposed configuration
  rc = lp_logits(P, C, Lx, Ly, B, pcfg, logits, precision, st);
  if (rc) return rc;
  hipLaunchKernelGGL(lp_choose_kernel, dim3(g), dim3(64 * LPC_WAVES), 0, st, B, N, Lx, Ly, E, 0,
                     logits, seed, step, (const double *)nullptr, pcfg, move, logq_r,
                     (int64_t *)nullptr, (const int64_t *)nullptr);
  VMCX_LAUNCHED();
  return 0;
}

// The same step with the logits at the current configurations given
// (logits_x, [B][3E]): only the proposed configurations go through the net.
// The chain is the one of lp_propose: logits at x are those the net gave
// when x was proposed (or the first evaluation), the same bits.
int lp_propose_given(const float *P, int C, int Lx, int Ly, int64_t B, const int64_t *cfg,
                     const double *logits_x, uint64_t seed, int64_t step,
                     const double *u_choice, double *logits_p, int32_t *move, double *logq_f,
                     int64_t *pcfg, double *logq_r, int precision, hipStream_t st,
                     const int64_t *step_dev) {
  if (lp_check(C, Lx, Ly)) return 1;
  VMCX_CHECK_ARG(B >= 0, "B < 0");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(P && cfg && logits_x && logits_p && move && logq_f && pcfg && logq_r,
                 "null pointer");
  const int N = Lx * Ly, E = lp_n_edges(Lx, Ly);
  const unsigned g = (unsigned)((B + LPC_WAVES - 1) / LPC_WAVES);
  hipLaunchKernelGGL(lp_choose_kernel, dim3(g), dim3(64 * LPC_WAVES), 0, st, B, N, Lx, Ly, E, 1,
                     logits_x, seed, step, u_choice, cfg, move, logq_f, pcfg, step_dev);
  VMCX_LAUNCHED();
  int rc = lp_logits(P, C, Lx, Ly, B, pcfg, logits_p, precision, st);
  if (rc) return rc;
  hipLaunchKernelGGL(lp_choose_kernel, dim3(g), dim3(64 * LPC_WAVES), 0, st, B, N, Lx, Ly, E, 0,
                     logits_p, seed, step, (const double *)nullptr, pcfg, move, logq_r,
                     (int64_t *)nullptr, step_dev);
  VMCX_LAUNCHED();
  return 0;
}

// logits_x[b] = logits_p[b] for the walkers whose move was accepted
__global__ void lp_commit_logits_kernel(int64_t B, int M, const uint8_t *__restrict__ accepted,
                                        const double *__restrict__ logits_p, double *logits_x) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * M) return;
  if (accepted[idx / M]) logits_x[idx] = logits_p[idx];
}

int lp_commit_logits(int64_t B, int M, const uint8_t *accepted, const double *logits_p,
                     double *logits_x, hipStream_t st) {
  VMCX_CHECK_ARG(B >= 0 && M >= 1, "bad sizes");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(accepted && logits_p && logits_x, "null pointer");
  const int64_t n = B * M;
  hipLaunchKernelGGL(lp_commit_logits_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     st, B, M, accepted, logits_p, logits_x);
  VMCX_LAUNCHED();
  return 0;
}

int lp_accept(int64_t B, int N, uint64_t seed, int64_t step, const double *u_acc,
              const int32_t *move, const int64_t *pcfg, const double *psign,
              const double *plog, const double *logq_f, const double *logq_r, int64_t *cfg,
              double *sign, double *logabs, int64_t *n_accept, uint8_t *accepted,
              hipStream_t st, const int64_t *step_dev) {
  VMCX_CHECK_ARG(B >= 0 && N >= 1 && N <= LP_NMAX, "bad sizes");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(move && pcfg && psign && plog && logq_f && logq_r && cfg && sign && logabs,
                 "null pointer");
  hipLaunchKernelGGL(lp_accept_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, st, B,
                     N, seed, step, u_acc, move, pcfg, psign, plog, logq_f, logq_r, cfg, sign,
                     logabs, n_accept, accepted, step_dev);
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
