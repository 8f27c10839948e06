// Neural-network backflow (NNBF) walkers on gfx950.
//
//   n   = occupation [up | down] (2N), one-hot input
//   z1  = W1 n + b1          (gather-sum of the W1 columns of occupied orbitals)
//   h1  = act(z1)
//   z2  = W2 h1 + b2         GEMM, batch of walkers / configurations
//   h2  = act(z2)
//   D   = Wout h2 + bout     GEMM  (2N*n_f outputs)
//   psi = det((M + s D)[occupied rows])   FP64 Gauss-Jordan in registers
//
// GEMMs run on MFMA through rocBLAS gemm_ex: FP32 operands, or BF16 operands
// with FP32 accumulation.  The sweep proposes one edge for all walkers at a
// time; z1 is carried per walker and updated by the W1 columns of the moved
// electrons, so only the two dense layers and the determinant are recomputed
// per proposal.
#include <hip/hip_bf16.h>
#include <rocblas/rocblas.h>

#include "../../include/vmcx.h"
#include "gj.h"
#include "kernels.h"

namespace vmcx {

__device__ __forceinline__ float act_f(float z, int act) {
  return act == 0 ? fmaxf(z, 0.0f) : tanhf(z);
}

template <typename T>
__device__ __forceinline__ T to_t(float v);
template <>
__device__ __forceinline__ float to_t<float>(float v) { return v; }
template <>
__device__ __forceinline__ __hip_bfloat16 to_t<__hip_bfloat16>(float v) {
  return __float2bfloat16(v);
}

// ------------------------------------------------------------ layer one
// z1[w][u] = b1[u] + sum of W1[u][orbital] over the occupied spin-orbitals
// (site ascending, up before down), h1 = act(z1).  A workgroup handles
// L1_RB configurations x L1_UT hidden units: the W1 tile sits in LDS with
// row stride 2N + 1 (odd, so the 64 lanes of a wave -- 64 consecutive units
// -- read 64 different banks), each configuration's occupation is two 64-bit
// masks from a wave ballot.
constexpr int L1_RB = 64, L1_UT = 64;

template <typename TI>
__global__ void __launch_bounds__(256)
    nnbf_layer1_kernel(int64_t Bc, int N, int H, int act,
                       const int64_t *__restrict__ cfg, const float *__restrict__ W1,
                       const float *__restrict__ b1, float *z1, TI *h1) {
  __shared__ float wt[L1_UT * (128 + 1)];
  __shared__ uint64_t mu[L1_RB], md[L1_RB];
  const int n_orb = 2 * N, ld = n_orb + 1;
  const int64_t r0 = (int64_t)blockIdx.x * L1_RB;
  const int u0 = blockIdx.y * L1_UT;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int idx = t; idx < L1_UT * n_orb; idx += blockDim.x) {
    const int uu = idx / n_orb, o = idx - uu * n_orb;
    wt[uu * ld + o] = (u0 + uu < H) ? W1[(int64_t)(u0 + uu) * n_orb + o] : 0.0f;
  }
  for (int r = wave; r < L1_RB; r += blockDim.x / 64) {
    const int64_t w = r0 + r;
    int64_t x = 0;
    if (w < Bc && lane < N) x = cfg[w * N + lane];
    const uint64_t bu = __ballot((x & 2) != 0), bd = __ballot((x & 1) != 0);
    if (lane == 0) {
      mu[r] = bu;
      md[r] = bd;
    }
  }
  __syncthreads();
  const int u = u0 + lane;
  if (u >= H) return;
  const float bias = b1[u];
  for (int r = wave; r < L1_RB; r += blockDim.x / 64) {
    const int64_t w = r0 + r;
    if (w >= Bc) break;
    const uint64_t bu = mu[r], bd = md[r];
    uint64_t all = bu | bd;
    float z = bias;
    while (all) {
      const int s = __ffsll((long long)all) - 1;
      all &= all - 1;
      if ((bu >> s) & 1) z += wt[lane * ld + s];
      if ((bd >> s) & 1) z += wt[lane * ld + N + s];
    }
    z1[w * H + u] = z;
    h1[w * H + u] = to_t<TI>(act_f(z, act));
  }
}

// z += b (in place), h = act(z)
template <typename TI>
__global__ void __launch_bounds__(256)
    bias_act_kernel(int64_t rows, int H, int act, float *z, const float *__restrict__ b,
                    TI *h) {
  const int64_t n = rows * H;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float v = z[i] + b[i % H];
    z[i] = v;
    h[i] = to_t<TI>(act_f(v, act));
  }
}

// ------------------------------------------------------------ determinant
// Builds A^T row-per-lane: a[j] = A[j][col] = M[orb_j][col] + s (D[orb_j*nf+col]
// + bout[...]) for orb_j the j-th occupied orbital (ascending), with the
// occupation masks given explicitly (up, dn).
template <int NFP>
__device__ __forceinline__ void nnbf_rows(double (&a)[NFP], uint64_t up, uint64_t dn, int N,
                                          int nf, int col, const float *__restrict__ M,
                                          const float *__restrict__ Drow,
                                          const float *__restrict__ bout, double s) {
  uint64_t mu = up, md = dn;
#pragma unroll
  for (int j = 0; j < NFP; ++j) {
    int orb = -1;
    if (mu) {
      orb = __ffsll((long long)mu) - 1;
      mu &= mu - 1;
    } else if (md) {
      orb = N + __ffsll((long long)md) - 1;
      md &= md - 1;
    }
    if (j < nf && col < nf) {
      if (orb >= 0) {
        const int64_t e = (int64_t)orb * nf + col;
        a[j] = (double)M[e] + s * ((double)Drow[e] + (double)bout[e]);
      } else {
        a[j] = 0.0;
      }
    } else {
      a[j] = (j == col) ? 1.0 : 0.0;
    }
  }
}

template <int NFP>
__device__ __forceinline__ void group_masks(const int64_t *cfg, int N, int col, uint64_t &up,
                                            uint64_t &dn) {
  uint64_t u = 0, d = 0;
  for (int s = col; s < N; s += NFP) {
    const int64_t x = cfg[s];
    if (x & 2) u |= 1ull << s;
    if (x & 1) d |= 1ull << s;
  }
#pragma unroll
  for (int o = NFP / 2; o > 0; o >>= 1) {
    u |= ((uint64_t)__shfl_xor((uint32_t)(u >> 32), o, 64) << 32) |
         __shfl_xor((uint32_t)u, o, 64);
    d |= ((uint64_t)__shfl_xor((uint32_t)(d >> 32), o, 64) << 32) |
         __shfl_xor((uint32_t)d, o, 64);
  }
  up = u;
  dn = d;
}

// INV: also write A^-1 (Gauss-Jordan); otherwise the determinant only (LU)
template <int NFP, bool INV>
__global__ void __launch_bounds__(256)
    nnbf_det_kernel(int64_t Bc, int N, int nf, const int64_t *__restrict__ cfg,
                    const float *__restrict__ D, const float *__restrict__ M,
                    const float *__restrict__ bout, double s, double *sign,
                    double *logabs, double *ainv) {
  constexpr int WPW = 64 / NFP;
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = lane / NFP, col = lane % NFP;
  const int wl = wave * WPW + grp;
  const int per_block = (blockDim.x / 64) * WPW;
  const int64_t w = (int64_t)blockIdx.x * per_block + wl;
  if (w >= Bc) return;
  double *prow = reinterpret_cast<double *>(smem) + wl * NFP;
  uint8_t *qbuf = reinterpret_cast<uint8_t *>(smem) + (size_t)per_block * NFP * 8 + wl * NFP;
  uint64_t up, dn;
  group_masks<NFP>(cfg + w * N, N, col, up, dn);
  const bool count_ok = (__popcll(up) + __popcll(dn)) == nf;
  double a[NFP];
  nnbf_rows<NFP>(a, up, dn, N, nf, col, M, D + w * (int64_t)(2 * N * nf), bout, s);
  const GJResult r = INV ? gj_invert<NFP>(a, col, grp * NFP, prow, qbuf)
                         : lu_logdet<NFP>(a, col, grp * NFP, prow);
  const bool singular = r.singular || !count_ok;
  if (col == 0) {
    sign[w] = singular ? 0.0 : ((r.parity & 1) ? -1.0 : 1.0);
    logabs[w] = singular ? (count_ok ? log(1e-45) : __builtin_nan(""))
                         : fmax(r.logdet, log(1e-45));
  }
  if (INV && r.my_step >= 0) {
    double *dst = ainv + ((size_t)w * NFP + r.my_step) * NFP;
#pragma unroll
    for (int c = 0; c < NFP; ++c) dst[qbuf[c]] = singular ? 0.0 : a[c];
  }
}

// ------------------------------------------------------------ sweep pieces
// prop[w] = flag | ni << 8 | nj << 12 | nmoves << 16 ; moves[w] = a1 b1 a2 b2 (bytes)
__global__ void __launch_bounds__(256)
    nnbf_propose_kernel(int64_t B, int N, int i, int j, int e, int64_t sweep,
                        uint64_t seed, double keep_rate, const double *u_choice,
                        const uint8_t *bits, const double *u_acc_in,
                        const int64_t *__restrict__ cfg, int32_t *prop, uint32_t *moves,
                        double *u_acc) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= B) return;
  const int si = (int)cfg[w * N + i], sj = (int)cfg[w * N + j];
  if (si == sj) {
    prop[w] = 0;
    return;
  }
  double uc, ua;
  int b0, b2;
  if (u_choice == nullptr) {
    const u32x4 o = philox4x32_10((uint32_t)e, (uint32_t)sweep, (uint32_t)w, 0u,
                                  (uint32_t)seed, (uint32_t)(seed >> 32));
    uc = u53(o.x, o.y);
    ua = u53(o.z, o.w);
    b0 = o.y & 1u;
    b2 = (o.y >> 1) & 1u;
  } else {
    uc = u_choice[w];
    ua = u_acc_in[w];
    b0 = bits[w] & 1;
    b2 = (bits[w] >> 1) & 1;
  }
  const int nmi = (si == 0) ? 0 : ((si == 3) ? 2 : 1);
  const int nmj = (sj == 0) ? 0 : ((sj == 3) ? 2 : 1);
  const int dnum = nmi > nmj ? nmi - nmj : nmj - nmi;
  int ni, nj;
  if (uc < keep_rate || dnum == 1) {
    ni = sj;
    nj = si;
  } else if (dnum == 0) {
    ni = b0 ? 3 : 0;
    nj = b0 ? 0 : 3;
  } else {
    ni = b2 ? 2 : 1;
    nj = b2 ? 1 : 2;
  }
  uint32_t mv = 0;
  int nm = 0;
  if ((si ^ ni) & 2) {
    const int from = (si & 2) ? i : j, to = (si & 2) ? j : i;
    mv |= (uint32_t)(from | (to << 8)) << (16 * nm);
    ++nm;
  }
  if ((si ^ ni) & 1) {
    const int from = (si & 1) ? i : j, to = (si & 1) ? j : i;
    mv |= (uint32_t)((N + from) | ((N + to) << 8)) << (16 * nm);
    ++nm;
  }
  prop[w] = 1 | (ni << 8) | (nj << 12) | (nm << 16);
  moves[w] = mv;
  u_acc[w] = ua;
}

// h1p = act(z1 + sum_moves (W1[:, to] - W1[:, from]))
template <typename TI>
__global__ void __launch_bounds__(256)
    nnbf_h1_proposed_kernel(int64_t B, int N, int H, int act, const float *__restrict__ W1,
                            const float *__restrict__ z1, const int32_t *__restrict__ prop,
                            const uint32_t *__restrict__ moves, TI *h1p) {
  const int64_t n = B * H;
  const int n_orb = 2 * N;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t w = idx / H;
    const int u = (int)(idx - w * H);
    float z = z1[idx];
    const int p = prop[w];
    if (p & 1) {
      const int nm = (p >> 16) & 3;
      const uint32_t mv = moves[w];
      for (int m = 0; m < nm; ++m) {
        const int from = (mv >> (16 * m)) & 0xFF, to = (mv >> (16 * m + 8)) & 0xFF;
        z += W1[(int64_t)u * n_orb + to] - W1[(int64_t)u * n_orb + from];
      }
    }
    h1p[idx] = to_t<TI>(act_f(z, act));
  }
}

// determinant of the proposed configuration, Metropolis test, commit
template <int NFP>
__global__ void __launch_bounds__(256)
    nnbf_accept_kernel(int64_t B, int N, int nf, int H, int i, int j, int64_t *cfg,
                       const int32_t *__restrict__ prop, const uint32_t *__restrict__ moves,
                       const double *__restrict__ u_acc, const float *__restrict__ D,
                       const float *__restrict__ M, const float *__restrict__ bout, double s,
                       const float *__restrict__ W1, float *z1, double *sign, double *logabs,
                       int64_t *n_accept) {
  constexpr int WPW = 64 / NFP;
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = lane / NFP, col = lane % NFP;
  const int wl = wave * WPW + grp;
  const int per_block = (blockDim.x / 64) * WPW;
  const int64_t w = (int64_t)blockIdx.x * per_block + wl;
  if (w >= B) return;
  const int p = prop[w];
  if (!(p & 1)) return;
  double *prow = reinterpret_cast<double *>(smem) + wl * NFP;
  const int ni = (p >> 8) & 0xF, nj = (p >> 12) & 0xF;
  uint64_t up, dn;
  group_masks<NFP>(cfg + w * N, N, col, up, dn);
  // proposed occupation
  up = (up & ~((1ull << i) | (1ull << j))) | ((uint64_t)((ni >> 1) & 1) << i) |
       ((uint64_t)((nj >> 1) & 1) << j);
  dn = (dn & ~((1ull << i) | (1ull << j))) | ((uint64_t)(ni & 1) << i) |
       ((uint64_t)(nj & 1) << j);
  double a[NFP];
  nnbf_rows<NFP>(a, up, dn, N, nf, col, M, D + w * (int64_t)(2 * N * nf), bout, s);
  const GJResult r = lu_logdet<NFP>(a, col, grp * NFP, prow);
  const double lp = r.singular ? log(1e-45) : fmax(r.logdet, log(1e-45));
  const double ratio = exp(2.0 * (lp - logabs[w]));
  if (!(u_acc[w] < ratio)) return;
  // commit (group-uniform branch)
  if (col == 0) {
    cfg[w * N + i] = ni;
    cfg[w * N + j] = nj;
    sign[w] = r.singular ? 0.0 : ((r.parity & 1) ? -1.0 : 1.0);
    logabs[w] = lp;
    if (n_accept) n_accept[w] += 1;
  }
  const int nm = (p >> 16) & 3;
  const uint32_t mv = moves[w];
  const int n_orb = 2 * N;
  for (int u = col; u < H; u += NFP) {
    float dz = 0.0f;
    for (int m = 0; m < nm; ++m) {
      const int from = (mv >> (16 * m)) & 0xFF, to = (mv >> (16 * m + 8)) & 0xFF;
      dz += W1[(int64_t)u * n_orb + to] - W1[(int64_t)u * n_orb + from];
    }
    z1[w * H + u] += dz;
  }
}

// ------------------------------------------------------------ local energy
// walker-major hop enumeration: counts, then etas / coefficients
__global__ void __launch_bounds__(256)
    conn_count_walker_kernel(int64_t B, int N, int E, const int32_t *__restrict__ edges,
                             const int64_t *__restrict__ cfg, int64_t *counts) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= B) return;
  const int64_t *c = cfg + w * N;
  int64_t k = 0;
  for (int e = 0; e < E; ++e) {
    const int i = edges[2 * e], j = edges[2 * e + 1];
    if ((unsigned)i >= (unsigned)N || (unsigned)j >= (unsigned)N || i == j) continue;
    k += (((c[i] >> 1) ^ (c[j] >> 1)) & 1) + ((c[i] ^ c[j]) & 1);
  }
  counts[w] = k;
}

// single-block exclusive scan (B up to millions; sequential chunks of 1024)
__global__ void __launch_bounds__(1024)
    exclusive_scan_kernel(int64_t n, const int64_t *in, int64_t *out, int64_t *total) {
  __shared__ int64_t sh[1024];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < n; base += blockDim.x) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = (i < n) ? in[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < (int)blockDim.x; o <<= 1) {
      const int64_t t = (threadIdx.x >= (unsigned)o) ? sh[threadIdx.x - o] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < n) out[i] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry += sh[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) total[0] = carry;
}

// bits a..b of a 64-bit mask (0 when a > b; 0 <= a, b <= 63)
__device__ __forceinline__ uint64_t bits_between(int a, int b) {
  if (a > b) return 0;
  const uint64_t hi = (b >= 63) ? ~0ull : ((1ull << (b + 1)) - 1);
  return hi & ~((1ull << a) - 1);
}

// One wave per walker (N <= 64: lane s holds site s): the 64 lanes take 32
// edges x 2 spins per round in (edge, spin) order, a ballot prefix gives each
// valid hop its slot, and every connected configuration is written by the
// 64 lanes at once (one site each).  The Jordan-Wigner parity of a hop i <-> j
// counts the modes strictly between them in the order (site, up before
// down): popcounts of the occupation masks over the site range.
__global__ void __launch_bounds__(256)
    conn_fill_walker_kernel(int64_t B, int N, int E, const int32_t *__restrict__ edges,
                            double t, const int64_t *__restrict__ cfg,
                            const int64_t *__restrict__ offsets, int64_t *etas,
                            double *coeffs) {
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= B) return;   // uniform per wave
  const int64_t cs = lane < N ? cfg[w * N + lane] : 0;
  const uint64_t up = __ballot(((cs >> 1) & 1) != 0), dn = __ballot((cs & 1) != 0);
  int64_t pos = offsets[w];
  for (int e0 = 0; e0 < E; e0 += 32) {
    const int e = e0 + (lane >> 1), sp = lane & 1;
    const int bit = sp == 0 ? 2 : 1;
    int i = 0, j = 0, par = 0;
    bool valid = false;
    if (e < E) {
      i = edges[2 * e];
      j = edges[2 * e + 1];
      if ((unsigned)i < (unsigned)N && (unsigned)j < (unsigned)N && i != j) {
        const uint64_t m = bit == 2 ? up : dn;
        valid = ((m >> i) & 1) != ((m >> j) & 1);
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        if (bit == 2)
          par = __popcll(up & bits_between(lo + 1, hi - 1)) +
                __popcll(dn & bits_between(lo + 1, hi));
        else
          par = __popcll(up & bits_between(lo, hi - 1)) +
                __popcll(dn & bits_between(lo + 1, hi - 1));
      }
    }
    const uint64_t vb = __ballot(valid);
    if (valid) coeffs[pos + __popcll(vb & ((1ull << lane) - 1))] = (par & 1) ? t : -t;
    uint64_t rem = vb;
    int64_t k = pos;
    while (rem) {
      const int src = __ffsll((long long)rem) - 1;
      rem &= rem - 1;
      const int ii = __shfl(i, src, 64), jj = __shfl(j, src, 64), bb = __shfl(bit, src, 64);
      if (lane < N) etas[k * N + lane] = (lane == ii || lane == jj) ? (cs ^ bb) : cs;
      ++k;
    }
    pos = k;
  }
}

int launch_conn_fill(int64_t B, int N, int E, const int32_t *edges, double t,
                     const int64_t *cfg, const int64_t *offsets, int64_t *etas, double *coeffs,
                     hipStream_t st) {
  VMCX_CHECK_ARG(N >= 1 && N <= 64, "connected-configuration fill needs 1 <= N <= 64");
  if (B == 0) return 0;
  hipLaunchKernelGGL(conn_fill_walker_kernel, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, st,
                     B, N, E, edges, t, cfg, offsets, etas, coeffs);
  VMCX_LAUNCHED();
  return 0;
}

__global__ void __launch_bounds__(256)
    eloc_segment_kernel(int64_t B, int N, double U, const int64_t *__restrict__ cfg,
                        const int64_t *__restrict__ offsets,
                        const int64_t *__restrict__ counts, const double *__restrict__ coeffs,
                        const double *__restrict__ csign, const double *__restrict__ clog,
                        const double *__restrict__ sign, const double *__restrict__ logabs,
                        double *eloc) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= B) return;
  const int64_t *c = cfg + w * N;
  int nd = 0;
  for (int s = 0; s < N; ++s) nd += (c[s] == 3);
  double acc = U * nd;
  const int64_t o = offsets[w], n = counts[w];
  for (int64_t k = 0; k < n; ++k)
    acc += coeffs[o + k] * csign[o + k] * sign[w] * exp(clog[o + k] - logabs[w]);
  eloc[w] = acc;
}

// ------------------------------------------------------------ backward
// O = d ln|psi| / d theta in the order [M, W1, b1, W2, b2, Wout, bout]:
//   gM = d ln|det A| / dA scattered to the occupied orbital rows (the M
//   block), dOut = s gM, dh2 = dOut Wout, dz2 = dh2 act'(z2), dh1 = dz2 W2,
//   dz1 = dh1 act'(z1); W1 = dz1 n^T, W2 = dz2 h1^T, Wout = dOut h2^T.
// The two products over the batch are SGEMMs; the per-sample kernels around
// them write the rows.  FACT: the row holds the factors instead (nD + 2N +
// 4H numbers): [gM | n | dz1 | h1 | dz2 | h2].

// head, one block per sample: gM (the M block of the row, TO) and s gM into
// G (FP32, [rows][nD], the GEMM operand)
template <typename TO>
__global__ void __launch_bounds__(256)
    nnbf_grad_head_kernel(int64_t Bc, int N, int nf, int nfp, const int64_t *__restrict__ cfg,
                          const double *__restrict__ ainv, double s, float *G, TO *O,
                          int64_t ld) {
  __shared__ int slot_of[128];
  const int64_t w = blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n_orb = 2 * N, nD = n_orb * nf;
  if (tid == 0) {
    int k = 0;
    const int64_t *c = cfg + w * N;
    for (int o = 0; o < n_orb; ++o) slot_of[o] = -1;
    for (int x = 0; x < N; ++x)
      if (c[x] & 2) slot_of[x] = k++;
    for (int x = 0; x < N; ++x)
      if (c[x] & 1) slot_of[N + x] = k++;
  }
  __syncthreads();
  const double *Ai = ainv + (size_t)w * nfp * nfp;
  const float sf = (float)s;
  TO *row = O + w * ld;
  float *g = G + w * (int64_t)nD;
  for (int idx = tid; idx < nD; idx += nt) {
    const int o = idx / nf, c = idx - o * nf;
    const int r = slot_of[o];
    // d ln|det A| / d A[r][c] = (A^-1)[c][r] = ainv[r][c]
    const float v = (r >= 0) ? (float)Ai[(size_t)r * nfp + c] : 0.0f;
    row[idx] = (TO)v;
    g[idx] = v * sf;
  }
}

// d *= act'(z) elementwise
__global__ void __launch_bounds__(256)
    nnbf_dact_kernel(int64_t n, int act, const float *__restrict__ z, float *d) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float v = z[i];
    d[i] = act == 0 ? (v > 0.0f ? d[i] : 0.0f) : d[i] * (1.0f - tanhf(v) * tanhf(v));
  }
}

// tail, one block per sample: the rest of the row from dz1, dz2 ([rows][H]),
// z1, z2 and s gM (G)
template <typename TO, bool FACT>
__global__ void __launch_bounds__(256)
    nnbf_grad_tail_kernel(int64_t Bc, int N, int nf, int H, int act,
                          const int64_t *__restrict__ cfg, const float *__restrict__ z1,
                          const float *__restrict__ z2, const float *__restrict__ dz1,
                          const float *__restrict__ dz2, const float *__restrict__ G, TO *O,
                          int64_t ld) {
  extern __shared__ float tsm[];
  float *h1 = tsm, *h2 = h1 + H, *d1 = h2 + H, *d2 = d1 + H, *nvec = d2 + H;   // 4H + 2N
  const int64_t w = blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n_orb = 2 * N, nD = n_orb * nf;
  const int64_t *c = cfg + w * N;
  for (int o = tid; o < n_orb; o += nt)
    nvec[o] = (o < N) ? ((c[o] & 2) ? 1.0f : 0.0f) : ((c[o - N] & 1) ? 1.0f : 0.0f);
  for (int u = tid; u < H; u += nt) {
    h1[u] = act_f(z1[w * H + u], act);
    h2[u] = act_f(z2[w * H + u], act);
    d1[u] = dz1[w * H + u];
    d2[u] = dz2[w * H + u];
  }
  __syncthreads();
  TO *row = O + w * ld;
  if (FACT) {
    TO *f = row + nD;
    for (int o = tid; o < n_orb; o += nt) f[o] = (TO)nvec[o];
    f += n_orb;
    for (int u = tid; u < H; u += nt) {
      f[u] = (TO)d1[u];
      f[H + u] = (TO)h1[u];
      f[2 * H + u] = (TO)d2[u];
      f[3 * H + u] = (TO)h2[u];
    }
    return;
  }
  const float *g = G + w * (int64_t)nD;
  int64_t off = nD;
  for (int64_t idx = tid; idx < (int64_t)H * n_orb; idx += nt) {   // W1
    const int u = (int)(idx / n_orb), jj = (int)(idx - (int64_t)u * n_orb);
    row[off + idx] = (TO)(d1[u] * nvec[jj]);
  }
  off += (int64_t)H * n_orb;
  for (int u = tid; u < H; u += nt) row[off + u] = (TO)d1[u];       // b1
  off += H;
  for (int64_t idx = tid; idx < (int64_t)H * H; idx += nt) {         // W2
    const int u = (int)(idx / H), i = (int)(idx - (int64_t)u * H);
    row[off + idx] = (TO)(d2[u] * h1[i]);
  }
  off += (int64_t)H * H;
  for (int u = tid; u < H; u += nt) row[off + u] = (TO)d2[u];       // b2
  off += H;
  for (int64_t idx = tid; idx < (int64_t)nD * H; idx += nt) {        // Wout
    const int i = (int)(idx / H), k = (int)(idx - (int64_t)i * H);
    row[off + idx] = (TO)(g[i] * h2[k]);
  }
  off += (int64_t)nD * H;
  for (int idx = tid; idx < nD; idx += nt) row[off + idx] = (TO)g[idx];   // bout
}

// ============================================================ host side
struct NnbfWs {
  // per-chunk forward buffers
  float *z1, *z2, *D;
  void *h1, *h2;
  float *g1, *g2;   // backward: dz1, dz2
  double *ainv;
  // sweep buffers (per walker)
  int32_t *prop;
  uint32_t *moves;
  double *uacc;
  // local-energy buffers
  int64_t *counts, *offsets, *total, *etas;
  double *coeffs, *csign, *clog;
  size_t bytes;
};

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

static NnbfWs nnbf_layout(const vmcx_nnbf_model &m, int64_t B, int64_t chunk, int n_edges,
                          char *base) {
  NnbfWs L{};
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t o = off;
    off += al(b);
    return base ? base + o : nullptr;
  };
  const int H = m.hidden, nD = 2 * m.n_sites * m.n_f, nfp = nfp_of(m.n_f);
  const size_t ti = m.gemm == 1 ? 2 : 4;
  L.z1 = (float *)take((size_t)chunk * H * 4);
  L.z2 = (float *)take((size_t)chunk * H * 4);
  L.D = (float *)take((size_t)chunk * nD * 4);
  L.h1 = take((size_t)chunk * H * ti);
  L.h2 = take((size_t)chunk * H * ti);
  L.g1 = (float *)take((size_t)chunk * H * 4);
  L.g2 = (float *)take((size_t)chunk * H * 4);
  L.ainv = (double *)take((size_t)chunk * nfp * nfp * 8);
  L.prop = (int32_t *)take((size_t)B * 4);
  L.moves = (uint32_t *)take((size_t)B * 4);
  L.uacc = (double *)take((size_t)B * 8);
  const int64_t Tmax = B * 2 * (int64_t)n_edges;
  L.counts = (int64_t *)take((size_t)B * 8);
  L.offsets = (int64_t *)take((size_t)B * 8);
  L.total = (int64_t *)take(8);
  L.etas = (int64_t *)take((size_t)Tmax * m.n_sites * 8);
  L.coeffs = (double *)take((size_t)Tmax * 8);
  L.csign = (double *)take((size_t)Tmax * 8);
  L.clog = (double *)take((size_t)Tmax * 8);
  L.bytes = off;
  return L;
}

static int check_model(const vmcx_nnbf_model *m) {
  VMCX_CHECK_ARG(m != nullptr, "null model");
  VMCX_CHECK_ARG(m->n_sites >= 1 && m->n_sites <= 64, "n_sites must be in [1, 64]");
  VMCX_CHECK_ARG(m->n_f >= 1 && m->n_f <= 64 && m->n_f <= 2 * m->n_sites, "bad n_f");
  VMCX_CHECK_ARG(m->hidden >= 1 && m->hidden <= 4096, "hidden must be in [1, 4096]");
  VMCX_CHECK_ARG(m->act == 0 || m->act == 1, "act must be 0 (ReLU) or 1 (tanh)");
  VMCX_CHECK_ARG(m->gemm == 0 || m->gemm == 1, "gemm must be 0 (FP32) or 1 (BF16)");
  VMCX_CHECK_ARG(m->M && m->W1 && m->b1 && m->W2 && m->b2 && m->Wout && m->bout,
                 "null weight pointer");
  VMCX_CHECK_ARG(m->gemm == 0 || (m->W2_bf16 && m->Wout_bf16), "BF16 mode needs BF16 weights");
  return 0;
}

int64_t nnbf_workspace_bytes(const vmcx_nnbf_model *m, int64_t B, int64_t chunk, int n_edges) {
  if (!m || B < 0 || chunk < 1 || n_edges < 0) return -1;
  return (int64_t)nnbf_layout(*m, B, chunk, n_edges, nullptr).bytes;
}

#define NN_BLAS(call)                                                         \
  do {                                                                       \
    rocblas_status s_ = (call);                                              \
    if (s_ != rocblas_status_success) {                                      \
      ::vmcx::set_error(std::string(__func__) + ": " #call " -> " +          \
                        rocblas_status_to_string(s_));                       \
      return 3;                                                              \
    }                                                                        \
  } while (0)

// z2 = W2 h1 ; D = Wout act(z2 + b2)  for `rows` rows already holding h1
static int dense_layers(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows,
                        hipStream_t st) {
  rocblas_handle h = blas_handle_for(st);
  VMCX_CHECK_ARG(h != nullptr, "rocblas handle");
  const int H = m.hidden, nD = 2 * m.n_sites * m.n_f;
  const float one = 1.0f, zero = 0.0f;
  const rocblas_datatype ti = m.gemm == 1 ? rocblas_datatype_bf16_r : rocblas_datatype_f32_r;
  const void *W2 = m.gemm == 1 ? m.W2_bf16 : (const void *)m.W2;
  const void *Wo = m.gemm == 1 ? m.Wout_bf16 : (const void *)m.Wout;
  // column-major view: z2 (H x rows) = W2 (H x H, row-major = W2^T col-major) * h1 (H x rows)
  NN_BLAS(rocblas_gemm_ex(h, rocblas_operation_transpose, rocblas_operation_none, H,
                          (rocblas_int)rows, H, &one, W2, ti, H, L.h1, ti, H, &zero, L.z2,
                          rocblas_datatype_f32_r, H, L.z2, rocblas_datatype_f32_r, H,
                          rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0));
  const unsigned g = (unsigned)(((rows * H) + 255) / 256 < 8192 ? ((rows * H) + 255) / 256 : 8192);
  if (m.gemm == 1)
    hipLaunchKernelGGL(bias_act_kernel<__hip_bfloat16>, dim3(g), dim3(256), 0, st, rows, H,
                       m.act, L.z2, m.b2, (__hip_bfloat16 *)L.h2);
  else
    hipLaunchKernelGGL(bias_act_kernel<float>, dim3(g), dim3(256), 0, st, rows, H, m.act,
                       L.z2, m.b2, (float *)L.h2);
  VMCX_LAUNCHED();
  NN_BLAS(rocblas_gemm_ex(h, rocblas_operation_transpose, rocblas_operation_none, nD,
                          (rocblas_int)rows, H, &one, Wo, ti, H, L.h2, ti, H, &zero, L.D,
                          rocblas_datatype_f32_r, nD, L.D, rocblas_datatype_f32_r, nD,
                          rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0));
  return 0;
}

template <int NFP>
static void launch_det(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows,
                       const int64_t *cfg, double *sign, double *logabs, double *ainv,
                       hipStream_t st) {
  constexpr int WPW = 64 / NFP;
  const int64_t per_block = 4 * WPW;
  const int64_t nblk = (rows + per_block - 1) / per_block;
  if (ainv)
    hipLaunchKernelGGL((nnbf_det_kernel<NFP, true>), dim3((unsigned)nblk), dim3(256),
                       (size_t)per_block * NFP * 9, st, rows, m.n_sites, m.n_f, cfg, L.D, m.M,
                       m.bout, m.bf_scale, sign, logabs, ainv);
  else
    hipLaunchKernelGGL((nnbf_det_kernel<NFP, false>), dim3((unsigned)nblk), dim3(256),
                       (size_t)per_block * NFP * 9, st, rows, m.n_sites, m.n_f, cfg, L.D, m.M,
                       m.bout, m.bf_scale, sign, logabs, ainv);
}

// forward of `rows` configurations (rows <= chunk); ainv optional
static int forward_rows(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows,
                        const int64_t *cfg, double *sign, double *logabs, double *ainv,
                        hipStream_t st) {
  if (rows == 0) return 0;
  const dim3 g1((unsigned)((rows + L1_RB - 1) / L1_RB),
                (unsigned)((m.hidden + L1_UT - 1) / L1_UT));
  if (m.gemm == 1)
    hipLaunchKernelGGL(nnbf_layer1_kernel<__hip_bfloat16>, g1, dim3(256), 0, st, rows,
                       m.n_sites, m.hidden, m.act, cfg, m.W1, m.b1, L.z1,
                       (__hip_bfloat16 *)L.h1);
  else
    hipLaunchKernelGGL(nnbf_layer1_kernel<float>, g1, dim3(256), 0, st, rows, m.n_sites,
                       m.hidden, m.act, cfg, m.W1, m.b1, L.z1, (float *)L.h1);
  VMCX_LAUNCHED();
  int rc = dense_layers(m, L, rows, st);
  if (rc) return rc;
  switch (nfp_of(m.n_f)) {
    case 8: launch_det<8>(m, L, rows, cfg, sign, logabs, ainv, st); break;
    case 16: launch_det<16>(m, L, rows, cfg, sign, logabs, ainv, st); break;
    case 32: launch_det<32>(m, L, rows, cfg, sign, logabs, ainv, st); break;
    default: launch_det<64>(m, L, rows, cfg, sign, logabs, ainv, st); break;
  }
  VMCX_LAUNCHED();
  return 0;
}

int nnbf_forward(const vmcx_nnbf_model *m, int64_t B, const int64_t *cfg, void *ws,
                 int64_t ws_bytes, int64_t chunk, double *sign, double *logabs, float *z1_out,
                 hipStream_t st) {
  if (check_model(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && chunk >= 1, "bad sizes");
  VMCX_CHECK_ARG(B == 0 || (cfg && sign && logabs && ws), "null pointer");
  const NnbfWs L = nnbf_layout(*m, B, chunk, 0, (char *)ws);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes, "workspace too small");
  for (int64_t r0 = 0; r0 < B; r0 += chunk) {
    const int64_t rows = (B - r0 < chunk) ? B - r0 : chunk;
    const int rc = forward_rows(*m, L, rows, cfg + r0 * m->n_sites, sign + r0, logabs + r0,
                                nullptr, st);
    if (rc) return rc;
    if (z1_out)
      VMCX_HIP(hipMemcpyAsync(z1_out + r0 * m->hidden, L.z1, (size_t)rows * m->hidden * 4,
                              hipMemcpyDeviceToDevice, st));
  }
  return 0;
}

template <int NFP>
static void launch_accept(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t B, int i, int j,
                          int64_t *cfg, float *z1, double *sign, double *logabs,
                          int64_t *n_accept, hipStream_t st) {
  constexpr int WPW = 64 / NFP;
  const int64_t per_block = 4 * WPW;
  const int64_t nblk = (B + per_block - 1) / per_block;
  hipLaunchKernelGGL(nnbf_accept_kernel<NFP>, dim3((unsigned)nblk), dim3(256),
                     (size_t)per_block * NFP * 9, st, B, m.n_sites, m.n_f, m.hidden, i, j, cfg,
                     L.prop, L.moves, L.uacc, L.D, m.M, m.bout, m.bf_scale, m.W1, z1, sign,
                     logabs, n_accept);
}

int nnbf_sweep(const vmcx_nnbf_model *m, int64_t B, int n_edges, const int32_t *edges_host,
               double hopping_rate, int n_sweeps, uint64_t seed, int64_t sweep_offset,
               const double *u_choice, const uint8_t *bits, const double *u_acc,
               int64_t *cfg, float *z1, double *sign, double *logabs, int64_t *n_accept,
               void *ws, int64_t ws_bytes, hipStream_t st) {
  if (check_model(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && n_edges >= 0 && n_sweeps >= 0, "negative size");
  VMCX_CHECK_ARG(B == 0 || (cfg && z1 && sign && logabs && ws), "null pointer");
  VMCX_CHECK_ARG(n_edges == 0 || edges_host, "null edges");
  VMCX_CHECK_ARG((u_choice == nullptr) == (u_acc == nullptr) &&
                     (u_choice == nullptr) == (bits == nullptr),
                 "injected draws need u_choice, bits and u_acc together");
  if (B == 0) return 0;
  const NnbfWs L = nnbf_layout(*m, B, B, 0, (char *)ws);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes, "workspace too small (chunk must be >= B)");
  const int N = m->n_sites;
  for (int e = 0; e < n_edges; ++e) {
    const int i = edges_host[2 * e], j = edges_host[2 * e + 1];
    VMCX_CHECK_ARG(i >= 0 && i < N && j >= 0 && j < N && i != j, "edge out of range");
  }
  const unsigned gw = (unsigned)((B + 255) / 256);
  const int64_t nh = B * m->hidden;
  const unsigned gh = (unsigned)((nh + 255) / 256 < 16384 ? (nh + 255) / 256 : 16384);
  for (int sw = 0; sw < n_sweeps; ++sw) {
    for (int e = 0; e < n_edges; ++e) {
      const int i = edges_host[2 * e], j = edges_host[2 * e + 1];
      const size_t di = ((size_t)sw * n_edges + e) * B;
      hipLaunchKernelGGL(nnbf_propose_kernel, dim3(gw), dim3(256), 0, st, B, N, i, j, e,
                         sweep_offset + sw, seed, 1.0 - hopping_rate,
                         u_choice ? u_choice + di : nullptr, bits ? bits + di : nullptr,
                         u_acc ? u_acc + di : nullptr, cfg, L.prop, L.moves, L.uacc);
      VMCX_LAUNCHED();
      if (m->gemm == 1)
        hipLaunchKernelGGL(nnbf_h1_proposed_kernel<__hip_bfloat16>, dim3(gh), dim3(256), 0, st,
                           B, N, m->hidden, m->act, m->W1, z1, L.prop, L.moves,
                           (__hip_bfloat16 *)L.h1);
      else
        hipLaunchKernelGGL(nnbf_h1_proposed_kernel<float>, dim3(gh), dim3(256), 0, st, B, N,
                           m->hidden, m->act, m->W1, z1, L.prop, L.moves, (float *)L.h1);
      VMCX_LAUNCHED();
      int rc = dense_layers(*m, L, B, st);
      if (rc) return rc;
      switch (nfp_of(m->n_f)) {
        case 8: launch_accept<8>(*m, L, B, i, j, cfg, z1, sign, logabs, n_accept, st); break;
        case 16: launch_accept<16>(*m, L, B, i, j, cfg, z1, sign, logabs, n_accept, st); break;
        case 32: launch_accept<32>(*m, L, B, i, j, cfg, z1, sign, logabs, n_accept, st); break;
        default: launch_accept<64>(*m, L, B, i, j, cfg, z1, sign, logabs, n_accept, st); break;
      }
      VMCX_LAUNCHED();
    }
  }
  return 0;
}

int nnbf_local_energy(const vmcx_nnbf_model *m, int64_t B, int n_edges, const int32_t *edges,
                      double t, double U, const int64_t *cfg, const double *sign,
                      const double *logabs, void *ws, int64_t ws_bytes, int64_t chunk,
                      double *eloc, hipStream_t st) {
  if (check_model(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && n_edges >= 0 && chunk >= 1, "bad sizes");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(cfg && sign && logabs && ws && eloc && (n_edges == 0 || edges),
                 "null pointer");
  const NnbfWs L = nnbf_layout(*m, B, chunk, n_edges, (char *)ws);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes, "workspace too small");
  const int N = m->n_sites;
  const unsigned gw = (unsigned)((B + 255) / 256);
  hipLaunchKernelGGL(conn_count_walker_kernel, dim3(gw), dim3(256), 0, st, B, N, n_edges,
                     edges, cfg, L.counts);
  VMCX_LAUNCHED();
  hipLaunchKernelGGL(exclusive_scan_kernel, dim3(1), dim3(1024), 0, st, B, L.counts,
                     L.offsets, L.total);
  VMCX_LAUNCHED();
  {
    const int rc = launch_conn_fill(B, N, n_edges, edges, t, cfg, L.offsets, L.etas, L.coeffs,
                                    st);
    if (rc) return rc;
  }
  int64_t T = 0;
  VMCX_HIP(hipMemcpyAsync(&T, L.total, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  VMCX_HIP(hipStreamSynchronize(st));
  for (int64_t r0 = 0; r0 < T; r0 += chunk) {
    const int64_t rows = (T - r0 < chunk) ? T - r0 : chunk;
    const int rc = forward_rows(*m, L, rows, L.etas + r0 * N, L.csign + r0, L.clog + r0,
                                nullptr, st);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(eloc_segment_kernel, dim3(gw), dim3(256), 0, st, B, N, U, cfg, L.offsets,
                     L.counts, L.coeffs, L.csign, L.clog, sign, logabs, eloc);
  VMCX_LAUNCHED();
  return 0;
}

// rows of O (or of its factors) for `rows` configurations whose forward
// pass left z1, z2 and A^-1 in the workspace
template <typename TO, bool FACT>
static int backward_rows(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows,
                         const int64_t *cfg, TO *O, int64_t ld, hipStream_t st) {
  const int N = m.n_sites, H = m.hidden, nf = m.n_f, nD = 2 * N * nf;
  rocblas_handle h = blas_handle_for(st);
  VMCX_CHECK_ARG(h != nullptr, "rocblas handle");
  hipLaunchKernelGGL(nnbf_grad_head_kernel<TO>, dim3((unsigned)rows), dim3(256), 0, st, rows, N,
                     nf, nfp_of(nf), cfg, L.ainv, m.bf_scale, L.D, O, ld);
  VMCX_LAUNCHED();
  const float one = 1.0f, zero = 0.0f;
  const int ri = (int)rows;
  // column-major views: dh2 (H x rows) = Wout^T (Wout row-major [nD][H]) * (s gM) (nD x rows)
  NN_BLAS(rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, H, ri, nD, &one,
                        m.Wout, H, L.D, nD, &zero, L.g2, H));
  const int64_t nh = rows * H;
  const unsigned gh = (unsigned)((nh + 255) / 256 < 8192 ? (nh + 255) / 256 : 8192);
  hipLaunchKernelGGL(nnbf_dact_kernel, dim3(gh), dim3(256), 0, st, nh, m.act, L.z2, L.g2);
  VMCX_LAUNCHED();
  // dh1 (H x rows) = W2^T-as-stored (W2 row-major [u][i]: column-major (i, u)) * dz2
  NN_BLAS(rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, H, ri, H, &one, m.W2,
                        H, L.g2, H, &zero, L.g1, H));
  hipLaunchKernelGGL(nnbf_dact_kernel, dim3(gh), dim3(256), 0, st, nh, m.act, L.z1, L.g1);
  VMCX_LAUNCHED();
  const size_t lds = (size_t)(4 * H + 2 * N) * sizeof(float);
  hipLaunchKernelGGL((nnbf_grad_tail_kernel<TO, FACT>), dim3((unsigned)rows), dim3(256), lds, st,
                     rows, N, nf, H, m.act, cfg, L.z1, L.z2, L.g1, L.g2, L.D, O, ld);
  VMCX_LAUNCHED();
  return 0;
}

int nnbf_logderiv(const vmcx_nnbf_model *m, int64_t B, const int64_t *cfg, void *ws,
                  int64_t ws_bytes, int64_t chunk, void *O, int o_dtype, int64_t ld,
                  double *sign, double *logabs, hipStream_t st) {
  if (check_model(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && chunk >= 1, "bad sizes");
  VMCX_CHECK_ARG(o_dtype == 0 || o_dtype == 1, "o_dtype must be 0 (f64) or 1 (f32)");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(cfg && ws && O && sign && logabs, "null pointer");
  const int N = m->n_sites, H = m->hidden, nf = m->n_f, nD = 2 * N * nf;
  const int64_t np = (int64_t)nD + (int64_t)H * 2 * N + H + (int64_t)H * H + H +
                     (int64_t)nD * H + nD;
  VMCX_CHECK_ARG(ld >= np, "ld smaller than the parameter count");
  const NnbfWs L = nnbf_layout(*m, B, chunk, 0, (char *)ws);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes, "workspace too small");
  VMCX_CHECK_ARG((4 * H + 2 * N) * 4 <= 65536, "backward LDS exceeds 64 KB");
  for (int64_t r0 = 0; r0 < B; r0 += chunk) {
    const int64_t rows = (B - r0 < chunk) ? B - r0 : chunk;
    int rc = forward_rows(*m, L, rows, cfg + r0 * N, sign + r0, logabs + r0, L.ainv, st);
    if (rc) return rc;
    rc = o_dtype == 0
             ? backward_rows<double, false>(*m, L, rows, cfg + r0 * N, (double *)O + r0 * ld,
                                            ld, st)
             : backward_rows<float, false>(*m, L, rows, cfg + r0 * N, (float *)O + r0 * ld, ld,
                                           st);
    if (rc) return rc;
  }
  return 0;
}

int nnbf_logderiv_factors(const vmcx_nnbf_model *m, int64_t B, const int64_t *cfg, void *ws,
                          int64_t ws_bytes, int64_t chunk, double *F, int64_t ld,
                          double *sign, double *logabs, hipStream_t st) {
  if (check_model(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && chunk >= 1, "bad sizes");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(cfg && ws && F && sign && logabs, "null pointer");
  const int N = m->n_sites, H = m->hidden, nf = m->n_f, nD = 2 * N * nf;
  VMCX_CHECK_ARG(ld >= (int64_t)nD + 2 * N + 4 * H, "ld smaller than the factor width");
  const NnbfWs L = nnbf_layout(*m, B, chunk, 0, (char *)ws);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes, "workspace too small");
  VMCX_CHECK_ARG((4 * H + 2 * N) * 4 <= 65536, "backward LDS exceeds 64 KB");
  for (int64_t r0 = 0; r0 < B; r0 += chunk) {
    const int64_t rows = (B - r0 < chunk) ? B - r0 : chunk;
    int rc = forward_rows(*m, L, rows, cfg + r0 * N, sign + r0, logabs + r0, L.ainv, st);
    if (rc) return rc;
    rc = backward_rows<double, true>(*m, L, rows, cfg + r0 * N, F + r0 * ld, ld, st);
    if (rc) return rc;
  }
  return 0;
}

// FP32 -> BF16 conversion of weight copies
__global__ void f32_to_bf16_kernel(int64_t n, const float *x, __hip_bfloat16 *y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = __float2bfloat16(x[i]);
}

int to_bf16(int64_t n, const float *x, void *y, hipStream_t st) {
  VMCX_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return 0;
  VMCX_CHECK_ARG(x && y, "null pointer");
  const unsigned g = (unsigned)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384);
  hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(g), dim3(256), 0, st, n, x, (__hip_bfloat16 *)y);
  VMCX_LAUNCHED();
  return 0;
}

}  // namespace vmcx
