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
// GEMMs run on the matrix cores through rocBLAS: FP32 operands (gemm_ex), BF16
// operands with FP32 accumulation (gemm_ex), or FP64 throughout (dgemm; the
// weights, activations and backflow in FP64, the plain FP64 form).  The
// kernels below take the scalar type TW of weights and activations (float for
// the FP32 and BF16 modes, double for FP64) and the GEMM input type TI.  The sweep proposes one edge for all walkers at a
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
__device__ __forceinline__ double act_f(double z, int act) {
  return act == 0 ? fmax(z, 0.0) : tanh(z);
}

template <typename T>
__device__ __forceinline__ T to_t(float v);
template <>
__device__ __forceinline__ float to_t<float>(float v) { return v; }
template <>
__device__ __forceinline__ __hip_bfloat16 to_t<__hip_bfloat16>(float v) {
  return __float2bfloat16(v);
}
template <typename T>
__device__ __forceinline__ T to_t(double v);
template <>
__device__ __forceinline__ double to_t<double>(double v) { return v; }

// ------------------------------------------------------------ layer one
// z1[w][u] = b1[u] + sum of W1[u][orbital] over the occupied spin-orbitals
// (site ascending, up before down), h1 = act(z1).  A workgroup handles
// L1_RB configurations x L1_UT hidden units: the W1 tile sits in LDS with
// row stride 2N + 1 (odd, so the 64 lanes of a wave -- 64 consecutive units
// -- read 64 different banks), each configuration's occupation is two 64-bit
// masks from a wave ballot.
constexpr int L1_RB = 64, L1_UT = 64;

// RB: configurations per workgroup (64; 16 for small batches, which then
// spread over four times the workgroups)
template <typename TW, typename TI, int RB = L1_RB>
__global__ void __launch_bounds__(256)
    nnbf_layer1_kernel(int64_t Bc, int N, int H, int act,
                       const int64_t *__restrict__ cfg, const TW *__restrict__ W1,
                       const TW *__restrict__ b1, TW *z1, TI *h1) {
  // dynamic LDS: the W1 tile [L1_UT][2N + 1] then the masks [2][RB]
  extern __shared__ __align__(16) char l1sm[];
  const int n_orb = 2 * N, ld = n_orb + 1;
  TW *wt = reinterpret_cast<TW *>(l1sm);
  uint64_t *mu = reinterpret_cast<uint64_t *>(l1sm + (((size_t)L1_UT * ld * sizeof(TW) + 15) &
                                                      ~(size_t)15));
  uint64_t *md = mu + RB;
  const int64_t r0 = (int64_t)blockIdx.x * RB;
  const int u0 = blockIdx.y * L1_UT;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int idx = t; idx < L1_UT * n_orb; idx += blockDim.x) {
    const int uu = idx / n_orb, o = idx - uu * n_orb;
    wt[uu * ld + o] = (u0 + uu < H) ? W1[(int64_t)(u0 + uu) * n_orb + o] : TW(0);
  }
  for (int r = wave; r < RB; r += blockDim.x / 64) {
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
  const TW bias = b1[u];
  for (int r = wave; r < RB; r += blockDim.x / 64) {
    const int64_t w = r0 + r;
    if (w >= Bc) break;
    const uint64_t bu = mu[r], bd = md[r];
    uint64_t all = bu | bd;
    TW z = bias;
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
template <typename TW, typename TI>
__global__ void __launch_bounds__(256)
    bias_act_kernel(int64_t rows, int H, int act, TW *z, const TW *__restrict__ b, TI *h) {
  const int64_t n = rows * H;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const TW v = z[i] + b[i % H];
    z[i] = v;
    h[i] = to_t<TI>(act_f(v, act));
  }
}

// ------------------------------------------------------------ determinant
// Builds A^T row-per-lane: a[j] = A[j][col] = M[orb_j][col] + s (D[orb_j*nf+col]
// + bout[...]) for orb_j the j-th occupied orbital (ascending), with the
// occupation masks given explicitly (up, dn).
template <int NFP, typename TW>
__device__ __forceinline__ void nnbf_rows(double (&a)[NFP], uint64_t up, uint64_t dn, int N,
                                          int nf, int col, const TW *__restrict__ M,
                                          const TW *__restrict__ Drow,
                                          const TW *__restrict__ bout, double s) {
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
template <int NFP, bool INV, typename TW>
__global__ void __launch_bounds__(256)
    nnbf_det_kernel(int64_t Bc, int N, int nf, const int64_t *__restrict__ cfg,
                    const TW *__restrict__ D, const TW *__restrict__ M,
                    const TW *__restrict__ bout, double s, double *sign,
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
  nnbf_rows<NFP, TW>(a, up, dn, N, nf, col, M, D + w * (int64_t)(2 * N * nf), bout, s);
  const bool nonfinite = group_nonfinite<NFP>(a, grp * NFP);
  const GJResult r = INV ? gj_invert<NFP>(a, col, grp * NFP, prow, qbuf)
                         : lu_logdet<NFP>(a, col, grp * NFP, prow);
  const bool singular = r.singular || !count_ok;
  if (col == 0) {
    sign[w] = nonfinite ? __builtin_nan("")
                        : (singular ? 0.0 : ((r.parity & 1) ? -1.0 : 1.0));
    logabs[w] = (nonfinite || !count_ok) ? __builtin_nan("")
                                         : (singular ? log(1e-45) : fmax(r.logdet, log(1e-45)));
  }
  if (INV && r.my_step >= 0) {
    double *dst = ainv + ((size_t)w * NFP + r.my_step) * NFP;
#pragma unroll
    for (int c = 0; c < NFP; ++c)
      dst[qbuf[c]] = nonfinite ? __builtin_nan("") : (singular ? 0.0 : a[c]);
  }
}

// ------------------------------------------------ BF16 matrix-core layers
// The two dense layers of the BF16 mode as hand-written matrix-core GEMMs
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation).  A workgroup (4 waves)
// owns NG_RB = 64 configurations: their BF16 input rows h ([64][K], padded
// rows of K + 8 so that the 16-byte fragment reads of 16 lanes fall in 64
// different banks) sit in LDS for the whole launch; each wave walks 32-row
// tiles of the weight matrix W ([n_rows][K] BF16, the reference's [out][in]
// layout), whose fragments it reads from global memory (L2), and keeps two
// 32 x 32 accumulator tiles (the 64 configurations).  MFMA operands, lane l:
// a = W[n0 + l % 32][k0 + 8 (l / 32) ... + 7], b = h[r0 + l % 32][k0 + 8 (l / 32)
// ... + 7]; result register q of lane l is row n0 + 8 (q / 4) + 4 (l / 32) +
// q % 4 of column r0 + l % 32.
// Epilogues:
//   NG_ACT:  out_h[r][n] = bf16(act(acc + b[n])) (the hidden layer, its
//            pre-activation kept in z[r][n] when z != NULL);
//   NG_SCAT: the output layer restricted to the occupied orbitals: for row
//            n = o nf + j of an orbital o occupied in configuration r,
//            Docc[r][slot_r(o)][j] = acc + b[n] (FP32, [nf][nf] per
//            configuration; slot_r(o) = rank of o among the occupied
//            orbitals).  The unoccupied half of the rows is computed and
//            dropped: the GEMM stays dense while the full backflow D never
//            reaches HBM.  The determinant kernel adds M in FP64.
typedef __bf16 ng_bf8 __attribute__((ext_vector_type(8)));
typedef float ng_f16v __attribute__((ext_vector_type(16)));
constexpr int NG_RB = 64, NG_T = 256;
constexpr int NG_ACT = 0, NG_SCAT = 1;

// LDS bytes of nnbf_gemm_mfma_kernel for input width K
__host__ __device__ inline size_t ng_lds_bytes(int K) {
  return (size_t)NG_RB * (K + 8) * 2 + (size_t)NG_RB * 128;
}

// Epilogue of one 32-row tile n0 of W against two 32-configuration tiles
// (acc0: configurations r0 + 0..31, acc1: r0 + 32..63); the slot table rows
// of these configurations start at srow.
// bias rows of a lane's accumulator registers: bv[q] = b[n0 + 8 (q / 4) +
// 4 kh + q % 4] (0 past n_rows), loaded before the tile's products so that
// the epilogue does not wait for them
__device__ __forceinline__ void ng_bias(const float *__restrict__ bias, int n0, int n_rows,
                                        int kh, float (&bv)[16]) {
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int nb = n0 + 8 * g + 4 * kh;
    if (nb + 3 < n_rows && (n_rows & 3) == 0) {
      const float4 v = *reinterpret_cast<const float4 *>(bias + nb);
      bv[4 * g + 0] = v.x;
      bv[4 * g + 1] = v.y;
      bv[4 * g + 2] = v.z;
      bv[4 * g + 3] = v.w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) bv[4 * g + e] = (nb + e < n_rows) ? bias[nb + e] : 0.0f;
    }
  }
}

template <int EPI>
__device__ __forceinline__ void ng_epilogue(int n0, const ng_f16v &acc0, const ng_f16v &acc1,
                                            int64_t r0, int srow, int lane, int64_t R,
                                            int n_rows, const float (&bv)[16], int act,
                                            float *z, __hip_bfloat16 *hout,
                                            const uint8_t *slot, int nf, float *A) {
  const int kh = lane >> 5;
  // the values are formed before any conditional store (a load first used
  // inside a branch makes the compiler wait for every outstanding memory
  // operation, earlier stores included, at each branch); the empty
  // statements keep them there
  float vs[2][16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    vs[0][q] = acc0[q] + bv[q];
    vs[1][q] = acc1[q] + bv[q];
    asm volatile("" : "+v"(vs[0][q]), "+v"(vs[1][q]));
  }
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int rl = half * 32 + (lane & 31);
    const int64_t r = r0 + rl;
    if (r >= R) continue;
    const float *vh = vs[half];
    if (EPI == NG_ACT) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int n = n0 + 8 * (q >> 2) + 4 * kh + (q & 3);
        if (n >= n_rows) continue;
        const float v = vh[q];
        if (z) z[r * n_rows + n] = v;
        hout[r * n_rows + n] = __float2bfloat16(act == 0 ? fmaxf(v, 0.0f) : tanhf(v));
      }
    } else {
      float *Ar = A + r * (int64_t)nf * nf;
      const uint8_t *srow_p = slot + (srow + rl) * 128;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int nb = n0 + 8 * g + 4 * kh;       // rows nb .. nb + 3
        if (nb >= n_rows) continue;
        const int o = nb / nf, j = nb - o * nf;
        const int sl = srow_p[o];
        if (sl == 0xFF) {
          // an orbital boundary inside the group (nf not a multiple of 4):
          // the other rows may belong to an occupied orbital
          if ((nf & 3) == 0 || j + 3 < nf) continue;
        }
        if ((nf & 3) == 0) {
          float4 v;
          v.x = vh[4 * g + 0];
          v.y = vh[4 * g + 1];
          v.z = vh[4 * g + 2];
          v.w = vh[4 * g + 3];
          *reinterpret_cast<float4 *>(Ar + (int64_t)sl * nf + j) = v;
        } else {
          for (int e = 0; e < 4; ++e) {
            const int n = nb + e;
            if (n >= n_rows) break;
            const int oo = n / nf, jj = n - oo * nf;
            const int ss = srow_p[oo];
            if (ss == 0xFF) continue;
            Ar[(int64_t)ss * nf + jj] = vh[4 * g + e];
          }
        }
      }
    }
  }
}

// slot table rows srow .. srow + nr - 1 for configurations c0 .. c0 + nr - 1
// (one wave): slot[(srow + r) * 128 + o] = rank of orbital o among the
// occupied orbitals of configuration c0 + r (0xFF: empty, or past nf)
__device__ __forceinline__ void ng_slot_row(int64_t x, int N, int nf, int lane, uint8_t *row) {
  const uint64_t up = __ballot((x & 2) != 0), dn = __ballot((x & 1) != 0);
  const int nup = __popcll(up);
  // entry e of the row: orbital e (up orbitals 0 .. N-1, down N .. 2N-1)
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int e = h * 64 + lane;
    uint8_t v = 0xFF;
    if (e < 2 * N) {
      const bool is_dn = e >= N;
      const int site = is_dn ? e - N : e;
      const uint64_t m = is_dn ? dn : up;
      const uint64_t below = (site == 0) ? 0ull : ((1ull << site) - 1ull);
      const int sl = (is_dn ? nup : 0) + __popcll(m & below);
      // a slot past nf (more occupied orbitals than nf) is dropped: the
      // determinant kernel reports such a configuration as NaN
      if (((m >> site) & 1ull) && sl < nf) v = (uint8_t)sl;
    }
    row[e] = v;
  }
}

__device__ __forceinline__ void ng_slots(const int64_t *__restrict__ cfg, int64_t c0, int nr,
                                         int64_t R, int N, int nf, int lane, uint8_t *slot,
                                         int srow) {
  // eight configurations' loads in flight at a time
  int r = 0;
  for (; r + 8 <= nr; r += 8) {
    int64_t x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      x[u] = (c0 + r + u < R && lane < N) ? cfg[(c0 + r + u) * N + lane] : 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) ng_slot_row(x[u], N, nf, lane, slot + (srow + r + u) * 128);
  }
  for (; r < nr; ++r) {
    int64_t x = 0;
    if (c0 + r < R && lane < N) x = cfg[(c0 + r) * N + lane];
    ng_slot_row(x, N, nf, lane, slot + (srow + r) * 128);
  }
}

// KFIX > 0: the input width known at compile time (the k loop fully unrolled,
// all weight-fragment loads of a tile in flight at once); 0: K at run time.
template <int EPI, int KFIX>
__global__ void __launch_bounds__(NG_T)
    nnbf_gemm_mfma_kernel(int64_t R, int n_rows, int Kr, const __hip_bfloat16 *__restrict__ W,
                          const __hip_bfloat16 *__restrict__ hin, const float *__restrict__ bias,
                          int act, float *z, __hip_bfloat16 *hout,
                          const int64_t *__restrict__ cfg, int N, int nf, float *A) {
  extern __shared__ __align__(16) char ngsm[];
  const int K = KFIX > 0 ? KFIX : Kr;
  const int ldh = K + 8;
  __bf16 *hs = reinterpret_cast<__bf16 *>(ngsm);
  uint8_t *slot = reinterpret_cast<uint8_t *>(ngsm + (size_t)NG_RB * ldh * 2);  // [64][128]
  const int64_t r0 = (int64_t)blockIdx.x * NG_RB;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, kh = lane >> 5;
  // stage the 64 input rows (16-byte chunks; rows past R are zero)
  const int kc = K / 8;
  for (int idx = t; idx < NG_RB * kc; idx += NG_T) {
    const int r = idx / kc, c = idx - r * kc;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (r0 + r < R) v = reinterpret_cast<const uint4 *>(hin + (r0 + r) * K)[c];
    *reinterpret_cast<uint4 *>(hs + r * ldh + c * 8) = v;
  }
  if (EPI == NG_SCAT)
    for (int r = wave; r < NG_RB; r += NG_T / 64) ng_slots(cfg, r0 + r, 1, R, N, nf, lane, slot, r);
  __syncthreads();
  // blockIdx.y: a contiguous slice of the 32-row tiles of W (small batches
  // spread over more workgroups)
  const int n_tiles = (n_rows + 31) / 32;
  const int tps = (n_tiles + (int)gridDim.y - 1) / (int)gridDim.y;
  const int t_end = min(n_tiles, ((int)blockIdx.y + 1) * tps);
  for (int tile = (int)blockIdx.y * tps + wave; tile < t_end; tile += NG_T / 64) {
    const int n0 = tile * 32;
    const int nrow = n0 + (lane & 31);
    const __bf16 *wrow = reinterpret_cast<const __bf16 *>(W) +
                         (int64_t)(nrow < n_rows ? nrow : n_rows - 1) * K + 8 * kh;
    ng_f16v acc0, acc1;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      acc0[q] = 0.0f;
      acc1[q] = 0.0f;
    }
    float bv[16];
    ng_bias(bias, n0, n_rows, kh, bv);
    const __bf16 *b0p = hs + (lane & 31) * ldh + 8 * kh;
    const __bf16 *b1p = hs + (32 + (lane & 31)) * ldh + 8 * kh;
#pragma unroll (KFIX > 0 ? KFIX / 16 : 4)
    for (int k0 = 0; k0 < K; k0 += 16) {
      const ng_bf8 a = *reinterpret_cast<const ng_bf8 *>(wrow + k0);
      const ng_bf8 b0 = *reinterpret_cast<const ng_bf8 *>(b0p + k0);
      const ng_bf8 b1 = *reinterpret_cast<const ng_bf8 *>(b1p + k0);
      acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b0, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b1, acc1, 0, 0, 0);
    }
    ng_epilogue<EPI>(n0, acc0, acc1, r0, 0, lane, R, n_rows, bv, act, z, hout, slot, nf, A);
  }
}

// The output layer for input width 256 (NG_SCAT only), organised around
// the configurations: a wave keeps the fragments of its 64 configurations
// (h rows, all 16 k-steps: 128 registers) for the whole launch, and the
// workgroup's four waves stream W's 32-row tiles through LDS,
// double-buffered: while a tile is multiplied, the next one (16 KiB,
// contiguous in the row-major W) is loaded by all 256 threads into
// registers and written to the other buffer after the tile's epilogue.
// Each W fragment read from LDS feeds two MFMAs (the wave's two
// configuration tiles), each W byte fetched from L2 serves 256
// configurations.  W tile rows are padded to 264 BF16 in LDS (16-byte
// fragment reads of 16 lanes fall in distinct banks).
// The configurations are the MFMA's A operand and W^T its B operand, so
// that result register q of lane l holds configuration 8 (q / 4) + 4 (l /
// 32) + q % 4 of the tile at output row n0 + l % 32: a store instruction
// writes 32 consecutive outputs (one orbital row segment, 128 bytes when
// nf is a multiple of 32) of two configurations.  The slot table is
// orbital-major (slotT[o][configuration], rows of NG2_SLD bytes: the byte
// writes of its construction fall in distinct banks, and the four slots
// of configurations c .. c + 3 are one 32-bit read).
// Grid: (configurations / 256, slices of W's tiles).
constexpr int NG2_T = 256, NG2_CPW = 64, NG2_CPB = NG2_CPW * (NG2_T / 64);
constexpr int NG2_K = 256, NG2_LDW = NG2_K + 8, NG2_SLD = NG2_CPB + 4;

// slotT[o * NG2_SLD + srow + r] for configurations c0 .. c0 + nr - 1 (one
// wave), eight configurations' loads in flight at a time
__device__ __forceinline__ void ng2_slot_col(int64_t x, int N, int nf, int lane, uint8_t *col) {
  const uint64_t up = __ballot((x & 2) != 0), dn = __ballot((x & 1) != 0);
  const int nup = __popcll(up);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int e = h * 64 + lane;
    uint8_t v = 0xFF;
    if (e < 2 * N) {
      const bool is_dn = e >= N;
      const int site = is_dn ? e - N : e;
      const uint64_t m = is_dn ? dn : up;
      const uint64_t below = (site == 0) ? 0ull : ((1ull << site) - 1ull);
      const int sl = (is_dn ? nup : 0) + __popcll(m & below);
      if (((m >> site) & 1ull) && sl < nf) v = (uint8_t)sl;
    }
    if (e < 128) col[e * NG2_SLD] = v;
  }
}

__device__ __forceinline__ void ng2_slots(const int64_t *__restrict__ cfg, int64_t c0, int nr,
                                          int64_t R, int N, int nf, int lane, uint8_t *slotT,
                                          int srow) {
  int r = 0;
  for (; r + 8 <= nr; r += 8) {
    int64_t x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      x[u] = (c0 + r + u < R && lane < N) ? cfg[(c0 + r + u) * N + lane] : 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) ng2_slot_col(x[u], N, nf, lane, slotT + srow + r + u);
  }
  for (; r < nr; ++r) {
    const int64_t x = (c0 + r < R && lane < N) ? cfg[(c0 + r) * N + lane] : 0;
    ng2_slot_col(x, N, nf, lane, slotT + srow + r);
  }
}

__global__ void __launch_bounds__(NG2_T)
    nnbf_ng2_kernel(int64_t R, int n_rows, const __hip_bfloat16 *__restrict__ W,
                    const __hip_bfloat16 *__restrict__ hin, const float *__restrict__ bias,
                    const int64_t *__restrict__ cfg, int N, int nf, float *A) {
  __shared__ __align__(16) __bf16 Ws[2 * 32 * NG2_LDW];     // [2][32][264]
  __shared__ __align__(16) uint8_t slotT[128 * NG2_SLD];     // [orbital][configuration]
  const int t = threadIdx.x, lane = t & 63, kh = lane >> 5;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);   // wave-uniform
  const int64_t rw = (int64_t)blockIdx.x * NG2_CPB + (int64_t)wave * NG2_CPW;
  const __bf16 *Wb = reinterpret_cast<const __bf16 *>(W);
  const __bf16 *hb = reinterpret_cast<const __bf16 *>(hin);
  // the wave's configuration fragments: configuration rw + 32 c + lane % 32,
  // k = 16 ks + 8 kh .. + 7
  ng_bf8 b[2][16];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int64_t row = rw + 32 * c + (lane & 31);
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
      ng_bf8 v;
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = (__bf16)0.0f;
      if (row < R) v = *reinterpret_cast<const ng_bf8 *>(hb + row * NG2_K + 16 * ks + 8 * kh);
      b[c][ks] = v;
    }
  }
  ng2_slots(cfg, rw, NG2_CPW, R, N, nf, lane, slotT, wave * NG2_CPW);
  // this workgroup's slice of W's tiles
  const int n_tiles = (n_rows + 31) / 32;
  const int tps = (n_tiles + (int)gridDim.y - 1) / (int)gridDim.y;
  const int t0 = (int)blockIdx.y * tps;
  const int t1 = min(n_tiles, t0 + tps);
  // tile staging: 16-byte chunk c = t + 256 i (i < 4) of the 16 KiB tile is
  // row c / 32, columns 8 (c % 32) .. + 7
  uint4 g[4];
  auto load_tile = [&](int tile) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c = t + NG2_T * i;
      const int row = tile * 32 + (c >> 5);
      g[i] = row < n_rows
                 ? *reinterpret_cast<const uint4 *>(Wb + (int64_t)row * NG2_K + 8 * (c & 31))
                 : make_uint4(0u, 0u, 0u, 0u);
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c = t + NG2_T * i;
      *reinterpret_cast<uint4 *>(Ws + (buf * 32 + (c >> 5)) * NG2_LDW + 8 * (c & 31)) = g[i];
    }
  };
  if (t0 < t1) {
    load_tile(t0);
    store_tile(0);
  }
  __syncthreads();
  // configurations of this lane's accumulator rows: rw + 32 half + 8 gq + 4 kh
  // + e; rows past R are not stored (lim: how many of the wave's rows from
  // 4 kh on exist)
  const int64_t lim64 = R - rw - 4 * kh;
  const int lim = lim64 > NG2_CPW ? NG2_CPW : (int)lim64;
  for (int tile = t0; tile < t1; ++tile) {
    const int buf = (tile - t0) & 1;
    const bool more = tile + 1 < t1;
    if (more) load_tile(tile + 1);
    // this lane's output row n of W: orbital o, column j of Docc
    const int n = tile * 32 + (lane & 31);
    const bool nok = n < n_rows;
    const int o = nok ? n / nf : 0, j = n - o * nf;
    const float bv = nok ? bias[n] : 0.0f;
    ng_f16v acc0, acc1;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      acc0[q] = 0.0f;
      acc1[q] = 0.0f;
    }
    const __bf16 *ap = Ws + (buf * 32 + (lane & 31)) * NG2_LDW + 8 * kh;
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
      const ng_bf8 a = *reinterpret_cast<const ng_bf8 *>(ap + 16 * ks);
      acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[0][ks], a, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[1][ks], a, acc1, 0, 0, 0);
    }
    // epilogue: Docc[r][slot_r(o)][j] = acc + b[n] for the configurations r
    // with orbital o occupied.  The row strides are made opaque here so that
    // the 32 row addresses are formed per store (scalar multiplies) instead
    // of being hoisted out of the tile loop into 64 registers.
    int64_t nf2 = (int64_t)nf * nf, rwt = rw;
    asm volatile("" : "+s"(nf2), "+s"(rwt));
    // The values and the slot words are formed before any conditional
    // store: a load whose result is first used inside a branch makes the
    // compiler wait for every outstanding memory operation (the stores
    // before it included) at each branch.
    float *Ab = A + (rwt + 4 * kh) * nf2 + j;
    const uint8_t *scol = slotT + o * NG2_SLD + wave * NG2_CPW + 4 * kh;
    float v[32];
    uint32_t w4[8];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      v[q] = acc0[q] + bv;
      v[16 + q] = acc1[q] + bv;
    }
#pragma unroll
    for (int gq = 0; gq < 8; ++gq)
      w4[gq] = *reinterpret_cast<const uint32_t *>(scol + 8 * gq);   // rows 8 gq + 4 kh ..
    // (empty statements that take the values as register operands keep
    // their computation, and its waits, ahead of the branches)
#pragma unroll
    for (int q = 0; q < 32; ++q) asm volatile("" : "+v"(v[q]));
#pragma unroll
    for (int gq = 0; gq < 8; ++gq) asm volatile("" : "+v"(w4[gq]));
#pragma unroll
    for (int gq = 0; gq < 8; ++gq) {
      const int rl0 = 8 * gq;                           // + 4 kh + e
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int sl = (int)((w4[gq] >> (8 * e)) & 0xFFu);
        if (nok && sl != 0xFF && rl0 + e < lim)
          Ab[(int64_t)(rl0 + e) * nf2 + sl * nf] = v[4 * gq + e];
      }
    }
    if (more) store_tile(buf ^ 1);
    __syncthreads();
  }
}

// Determinant from the compact Docc ([nf][nf] FP32 per configuration,
// written by the NG_SCAT epilogue): lane col of the group holds column col,
// a[j] = M[orb_j][col] + s Docc[j][col] for orb_j the j-th occupied orbital
// (ascending); padding as nnbf_rows.
// TL: the elimination's type (double; float for the FP32 determinant of the
// mixed-precision amplitude, whose entries are formed by one FP32 fma)
template <int NFP, typename TL>
__global__ void __launch_bounds__(256)
    nnbf_detA_kernel(int64_t Bc, int N, int nf, const int64_t *__restrict__ cfg,
                     const float *__restrict__ A, const float *__restrict__ M, double s,
                     double *sign, double *logabs) {
  constexpr int WPW = 64 / NFP;
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = lane / NFP, col = lane % NFP;
  const int wl = wave * WPW + grp;
  const int per_block = (blockDim.x / 64) * WPW;
  const int64_t w = (int64_t)blockIdx.x * per_block + wl;
  if (w >= Bc) return;
  TL *prow = reinterpret_cast<TL *>(smem) + wl * NFP;
  uint64_t up, dn;
  if constexpr (NFP == 64) {
    // one matrix per wave (N <= 64): the occupation masks by two ballots,
    // so that they and the orbital of every row stay in scalar registers
    const int64_t x = col < N ? cfg[w * N + col] : 0;
    up = __ballot((x & 2) != 0);
    dn = __ballot((x & 1) != 0);
  } else {
    group_masks<NFP>(cfg + w * N, N, col, up, dn);
  }
  const bool count_ok = (__popcll(up) + __popcll(dn)) == nf;
  const float *Aw = A + w * (int64_t)nf * nf;
  // FP32 elimination: the entries as fmaf(s, Docc, M), one rounding
  const float sf = (float)s;
  TL a[NFP];
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
      if (orb < 0)
        a[j] = TL(0);
      else if constexpr (sizeof(TL) == 4)
        a[j] = fmaf(sf, Aw[(int64_t)j * nf + col], M[(int64_t)orb * nf + col]);
      else
        a[j] = (double)M[(int64_t)orb * nf + col] + s * (double)Aw[(int64_t)j * nf + col];
    } else {
      a[j] = (j == col) ? TL(1) : TL(0);
    }
  }
  const bool nonfinite = group_nonfinite<NFP>(a, grp * NFP);
  const GJResult r = lu_logdet<NFP>(a, col, grp * NFP, prow);
  const bool singular = r.singular || !count_ok;
  if (col == 0) {
    sign[w] = nonfinite ? __builtin_nan("")
                        : (singular ? 0.0 : ((r.parity & 1) ? -1.0 : 1.0));
    logabs[w] = (nonfinite || !count_ok) ? __builtin_nan("")
                                         : (singular ? log(1e-45) : fmax(r.logdet, log(1e-45)));
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
template <typename TW, typename TI>
__global__ void __launch_bounds__(256)
    nnbf_h1_proposed_kernel(int64_t B, int N, int H, int act, const TW *__restrict__ W1,
                            const TW *__restrict__ z1, const int32_t *__restrict__ prop,
                            const uint32_t *__restrict__ moves, TI *h1p) {
  const int64_t n = B * H;
  const int n_orb = 2 * N;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t w = idx / H;
    const int u = (int)(idx - w * H);
    TW z = z1[idx];
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

// W1T[o][u] = W1[u][o]: the first-layer columns contiguous, so that the
// incremental kernels below read a column with coalesced loads
template <typename TW>
__global__ void __launch_bounds__(256)
    nnbf_w1t_kernel(int H, int n_orb, const TW *__restrict__ W1, TW *W1T) {
  const int64_t n = (int64_t)H * n_orb;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int o = (int)(idx / H), u = (int)(idx - (int64_t)o * H);
    W1T[idx] = W1[(int64_t)u * n_orb + o];
  }
}

// First layer of connected configurations by one hop each (local energy):
// h1[k] = act(z1w[parent_k] + W1[:, to_k] - W1[:, from_k]), the parent
// walker's pre-activations updated by the moved electron's column pair.
template <typename TW, typename TI>
__global__ void __launch_bounds__(256)
    nnbf_h1_conn_kernel(int64_t rows, int H, const TW *__restrict__ W1T,
                        const TW *__restrict__ z1w, const int32_t *__restrict__ cparent,
                        const uint32_t *__restrict__ cmove, int act, TI *h1) {
  const int64_t n = rows * H;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = idx / H;
    const int u = (int)(idx - k * H);
    const uint32_t mv = cmove[k];
    const int from = mv & 0xFF, to = (mv >> 8) & 0xFF;
    const TW z = z1w[(int64_t)cparent[k] * H + u] + W1T[(int64_t)to * H + u] -
                 W1T[(int64_t)from * H + u];
    h1[idx] = to_t<TI>(act_f(z, act));
  }
}

// First layer of proposed configurations (one wave per configuration):
// z1p[r] = z1[r] + sum over the orbitals o whose occupation differs between
// cfg[r] and pcfg[r] of (+1 if o is occupied in pcfg, else -1) W1[:, o]
// (orbitals in ascending order, up then down), h1[r] = act(z1p[r]).
template <typename TW, typename TI>
__global__ void __launch_bounds__(256)
    nnbf_h1_moved_kernel(int64_t rows, int N, int H, const TW *__restrict__ W1T,
                         const TW *__restrict__ z1, const int64_t *__restrict__ cfg,
                         const int64_t *__restrict__ pcfg, int act, TW *z1p, TI *h1) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (r >= rows) return;
  int64_t x = 0, y = 0;
  if (lane < N) {
    x = cfg[r * N + lane];
    y = pcfg[r * N + lane];
  }
  const uint64_t dup = __ballot(((x ^ y) & 2) != 0), ddn = __ballot(((x ^ y) & 1) != 0);
  const uint64_t nup = __ballot((y & 2) != 0), ndn = __ballot((y & 1) != 0);
  for (int u = lane; u < H; u += 64) {
    TW z = z1[r * H + u];
    for (uint64_t m = dup; m; m &= m - 1) {
      const int s = __ffsll((long long)m) - 1;
      const TW w = W1T[(int64_t)s * H + u];
      z = ((nup >> s) & 1ull) ? z + w : z - w;
    }
    for (uint64_t m = ddn; m; m &= m - 1) {
      const int s = __ffsll((long long)m) - 1;
      const TW w = W1T[(int64_t)(N + s) * H + u];
      z = ((ndn >> s) & 1ull) ? z + w : z - w;
    }
    z1p[r * H + u] = z;
    h1[r * H + u] = to_t<TI>(act_f(z, act));
  }
}

// rows of dst (row_words 32-bit words each) replaced by those of src where
// mask[r] != 0
__global__ void masked_rows_kernel(int64_t B, int64_t row_words, const uint8_t *__restrict__ mask,
                                   const uint32_t *__restrict__ src, uint32_t *dst) {
  const int64_t n = B * row_words;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x)
    if (mask[idx / row_words]) dst[idx] = src[idx];
}

// determinant of the proposed configuration, Metropolis test, commit
template <int NFP, typename TW>
__global__ void __launch_bounds__(256)
    nnbf_accept_kernel(int64_t B, int N, int nf, int H, int i, int j, int64_t *cfg,
                       const int32_t *__restrict__ prop, const uint32_t *__restrict__ moves,
                       const double *__restrict__ u_acc, const TW *__restrict__ D,
                       const TW *__restrict__ M, const TW *__restrict__ bout, double s,
                       const TW *__restrict__ W1, TW *z1, double *sign, double *logabs,
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
  nnbf_rows<NFP, TW>(a, up, dn, N, nf, col, M, D + w * (int64_t)(2 * N * nf), bout, s);
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
    TW dz = 0;
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
                            double *coeffs, int32_t *cparent, uint32_t *cmove) {
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
      if (cparent != nullptr && lane == 0) {
        const uint64_t mm = bb == 2 ? up : dn;
        const int from = ((mm >> ii) & 1ull) ? ii : jj, to = ii + jj - from;
        const int ob = bb == 2 ? 0 : N;
        cparent[k] = (int32_t)w;
        cmove[k] = (uint32_t)(ob + from) | ((uint32_t)(ob + to) << 8);
      }
      ++k;
    }
    pos = k;
  }
}

int launch_conn_fill(int64_t B, int N, int E, const int32_t *edges, double t,
                     const int64_t *cfg, const int64_t *offsets, int64_t *etas, double *coeffs,
                     hipStream_t st, int32_t *cparent, uint32_t *cmove) {
  VMCX_CHECK_ARG(N >= 1 && N <= 64, "connected-configuration fill needs 1 <= N <= 64");
  VMCX_CHECK_ARG((cparent == nullptr) == (cmove == nullptr), "cparent and cmove go together");
  if (B == 0) return 0;
  hipLaunchKernelGGL(conn_fill_walker_kernel, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, st,
                     B, N, E, edges, t, cfg, offsets, etas, coeffs, cparent, cmove);
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
template <typename TO, typename TW>
__global__ void __launch_bounds__(256)
    nnbf_grad_head_kernel(int64_t Bc, int N, int nf, int nfp, const int64_t *__restrict__ cfg,
                          const double *__restrict__ ainv, double s, TW *G, TO *O,
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
  const TW sf = (TW)s;
  TO *row = O + w * ld;
  TW *g = G + w * (int64_t)nD;
  for (int idx = tid; idx < nD; idx += nt) {
    const int o = idx / nf, c = idx - o * nf;
    const int r = slot_of[o];
    // d ln|det A| / d A[r][c] = (A^-1)[c][r] = ainv[r][c]
    const TW v = (r >= 0) ? (TW)Ai[(size_t)r * nfp + c] : TW(0);
    row[idx] = (TO)v;
    g[idx] = v * sf;
  }
}

// d *= act'(z) elementwise
template <typename TW>
__global__ void __launch_bounds__(256)
    nnbf_dact_kernel(int64_t n, int act, const TW *__restrict__ z, TW *d) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const TW v = z[i];
    if (act == 0) {
      d[i] = v > TW(0) ? d[i] : TW(0);
    } else {
      const TW th = act_f(v, 1);
      d[i] *= TW(1) - th * th;
    }
  }
}

// tail, one block per sample: the rest of the row from dz1, dz2 ([rows][H]),
// z1, z2 and s gM (G)
template <typename TO, bool FACT, typename TW>
__global__ void __launch_bounds__(256)
    nnbf_grad_tail_kernel(int64_t Bc, int N, int nf, int H, int act,
                          const int64_t *__restrict__ cfg, const TW *__restrict__ z1,
                          const TW *__restrict__ z2, const TW *__restrict__ dz1,
                          const TW *__restrict__ dz2, const TW *__restrict__ G, TO *O,
                          int64_t ld) {
  extern __shared__ __align__(16) char tsm_raw[];
  TW *tsm = reinterpret_cast<TW *>(tsm_raw);
  TW *h1 = tsm, *h2 = h1 + H, *d1 = h2 + H, *d2 = d1 + H, *nvec = d2 + H;   // 4H + 2N
  const int64_t w = blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n_orb = 2 * N, nD = n_orb * nf;
  const int64_t *c = cfg + w * N;
  for (int o = tid; o < n_orb; o += nt)
    nvec[o] = (o < N) ? ((c[o] & 2) ? TW(1) : TW(0)) : ((c[o - N] & 1) ? TW(1) : TW(0));
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
  const TW *g = G + w * (int64_t)nD;
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
// Weights of one mode: FP32 (gemm 0 / 1) or FP64 (gemm 2).
template <typename TW>
struct NnW {
  const TW *M, *W1, *b1, *W2, *b2, *Wout, *bout;
};
template <typename TW>
static NnW<TW> weights_of(const vmcx_nnbf_model &m);
template <>
NnW<float> weights_of<float>(const vmcx_nnbf_model &m) {
  return {m.M, m.W1, m.b1, m.W2, m.b2, m.Wout, m.bout};
}
template <>
NnW<double> weights_of<double>(const vmcx_nnbf_model &m) {
  return {m.M64, m.W1_64, m.b1_64, m.W2_64, m.b2_64, m.Wout_64, m.bout_64};
}

struct NnbfWs {
  // per-chunk forward buffers (elements of the mode's weight type TW, h1 / h2
  // of the GEMM input type)
  void *z1, *z2, *D;
  void *h1, *h2;
  void *g1, *g2;   // backward: dz1, dz2
  double *ainv;
  // sweep buffers (per walker)
  int32_t *prop;
  uint32_t *moves;
  double *uacc;
  // local-energy buffers
  int64_t *counts, *offsets, *total, *etas;
  double *coeffs, *csign, *clog;
  int32_t *cparent;   // walker of each connected configuration
  uint32_t *cmove;    // its hop: from_orbital | to_orbital << 8
  void *z1w;          // first-layer pre-activations of the walkers [B][H]
  void *w1t;          // W1 transposed [2N][H]
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
  const size_t tw = m.gemm == 2 ? 8 : 4;
  const size_t ti = m.gemm == 1 ? 2 : tw;
  L.z1 = take((size_t)chunk * H * tw);
  L.z2 = take((size_t)chunk * H * tw);
  L.D = take((size_t)chunk * nD * tw);
  L.h1 = take((size_t)chunk * H * ti);
  L.h2 = take((size_t)chunk * H * ti);
  L.g1 = take((size_t)chunk * H * tw);
  L.g2 = take((size_t)chunk * H * tw);
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
  L.cparent = (int32_t *)take((size_t)Tmax * 4);
  L.cmove = (uint32_t *)take((size_t)Tmax * 4);
  L.z1w = take((size_t)B * H * tw);
  L.w1t = take((size_t)2 * m.n_sites * H * tw);
  L.bytes = off;
  return L;
}

static int check_model(const vmcx_nnbf_model *m) {
  VMCX_CHECK_ARG(m != nullptr, "null model");
  VMCX_CHECK_ARG(m->n_sites >= 1 && m->n_sites <= 64, "n_sites must be in [1, 64]");
  VMCX_CHECK_ARG(m->n_f >= 1 && m->n_f <= 64 && m->n_f <= 2 * m->n_sites, "bad n_f");
  VMCX_CHECK_ARG(m->hidden >= 1 && m->hidden <= 4096, "hidden must be in [1, 4096]");
  VMCX_CHECK_ARG(m->act == 0 || m->act == 1, "act must be 0 (ReLU) or 1 (tanh)");
  VMCX_CHECK_ARG(m->gemm >= 0 && m->gemm <= 2, "gemm must be 0 (FP32), 1 (BF16) or 2 (FP64)");
  if (m->gemm == 2)
    VMCX_CHECK_ARG(m->M64 && m->W1_64 && m->b1_64 && m->W2_64 && m->b2_64 && m->Wout_64 &&
                       m->bout_64,
                   "FP64 mode needs the FP64 weight pointers");
  else
    VMCX_CHECK_ARG(m->M && m->W1 && m->b1 && m->W2 && m->b2 && m->Wout && m->bout,
                   "null weight pointer");
  VMCX_CHECK_ARG(m->gemm != 1 || (m->W2_bf16 && m->Wout_bf16), "BF16 mode needs BF16 weights");
  VMCX_CHECK_ARG(m->det_fp32 == 0 || (m->gemm == 1 && m->path == 0),
                 "det_fp32 needs the BF16 mode on the library's matrix-core layers");
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

// column-major C (m x n) = op(A) B for the dense layers: FP64 by dgemm, FP32 /
// BF16 operands with FP32 results by gemm_ex
static int nn_gemm(rocblas_handle h, rocblas_operation ta, int mm, int nn, int kk,
                   const void *A, int lda, const void *Bm, int ldb, void *C, int ldc, int gemm) {
  if (gemm == 2) {
    const double one = 1.0, zero = 0.0;
    NN_BLAS(rocblas_dgemm(h, ta, rocblas_operation_none, mm, nn, kk, &one, (const double *)A,
                          lda, (const double *)Bm, ldb, &zero, (double *)C, ldc));
    return 0;
  }
  const float one = 1.0f, zero = 0.0f;
  const rocblas_datatype ti = gemm == 1 ? rocblas_datatype_bf16_r : rocblas_datatype_f32_r;
  NN_BLAS(rocblas_gemm_ex(h, ta, rocblas_operation_none, mm, nn, kk, &one, A, ti, lda, Bm, ti,
                          ldb, &zero, C, rocblas_datatype_f32_r, ldc, C, rocblas_datatype_f32_r,
                          ldc, rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0));
  return 0;
}

// z2 = W2 h1 ; D = Wout act(z2 + b2)  for `rows` rows already holding h1
template <typename TW, typename TI>
static int dense_layers(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows,
                        hipStream_t st) {
  rocblas_handle h = blas_handle_for(st);
  VMCX_CHECK_ARG(h != nullptr, "rocblas handle");
  const NnW<TW> W = weights_of<TW>(m);
  const int H = m.hidden, nD = 2 * m.n_sites * m.n_f;
  const void *W2 = m.gemm == 1 ? m.W2_bf16 : (const void *)W.W2;
  const void *Wo = m.gemm == 1 ? m.Wout_bf16 : (const void *)W.Wout;
  // column-major view: z2 (H x rows) = W2 (H x H, row-major = W2^T col-major) * h1 (H x rows)
  int rc = nn_gemm(h, rocblas_operation_transpose, H, (int)rows, H, W2, H, L.h1, H, L.z2, H,
                   m.gemm);
  if (rc) return rc;
  const unsigned g = (unsigned)(((rows * H) + 255) / 256 < 8192 ? ((rows * H) + 255) / 256 : 8192);
  hipLaunchKernelGGL((bias_act_kernel<TW, TI>), dim3(g), dim3(256), 0, st, rows, H, m.act,
                     (TW *)L.z2, W.b2, (TI *)L.h2);
  VMCX_LAUNCHED();
  return nn_gemm(h, rocblas_operation_transpose, nD, (int)rows, H, Wo, H, L.h2, H, L.D, nD,
                 m.gemm);
}

template <int NFP, typename TW>
static void launch_det(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows,
                       const int64_t *cfg, double *sign, double *logabs, double *ainv,
                       hipStream_t st) {
  constexpr int WPW = 64 / NFP;
  const int64_t per_block = 4 * WPW;
  const int64_t nblk = (rows + per_block - 1) / per_block;
  const NnW<TW> W = weights_of<TW>(m);
  if (ainv)
    hipLaunchKernelGGL((nnbf_det_kernel<NFP, true, TW>), dim3((unsigned)nblk), dim3(256),
                       (size_t)per_block * NFP * 9, st, rows, m.n_sites, m.n_f, cfg,
                       (const TW *)L.D, W.M, W.bout, m.bf_scale, sign, logabs, ainv);
  else
    hipLaunchKernelGGL((nnbf_det_kernel<NFP, false, TW>), dim3((unsigned)nblk), dim3(256),
                       (size_t)per_block * NFP * 9, st, rows, m.n_sites, m.n_f, cfg,
                       (const TW *)L.D, W.M, W.bout, m.bf_scale, sign, logabs, ainv);
}

template <typename TW, typename TI>
static int layer1_to(const vmcx_nnbf_model &m, int64_t rows, const int64_t *cfg, TW *z1, TI *h1,
                     hipStream_t st) {
  const unsigned ub = (unsigned)((m.hidden + L1_UT - 1) / L1_UT);
  // small batches (fewer than ~1024 workgroups of 64 configurations) take
  // 16 configurations per workgroup
  const bool small = ((rows + L1_RB - 1) / L1_RB) * ub < 1024;
  const int rb = small ? 16 : L1_RB;
  const dim3 g1((unsigned)((rows + rb - 1) / rb), ub);
  const size_t lds = (((size_t)L1_UT * (2 * m.n_sites + 1) * sizeof(TW) + 15) & ~(size_t)15) +
                     2 * rb * sizeof(uint64_t);
  VMCX_CHECK_ARG(lds <= max_lds_per_block(), "layer-1 LDS exceeds the device limit");
  const NnW<TW> W = weights_of<TW>(m);
  if (small)
    hipLaunchKernelGGL((nnbf_layer1_kernel<TW, TI, 16>), g1, dim3(256), lds, st, rows,
                       m.n_sites, m.hidden, m.act, cfg, W.W1, W.b1, z1, h1);
  else
    hipLaunchKernelGGL((nnbf_layer1_kernel<TW, TI, L1_RB>), g1, dim3(256), lds, st, rows,
                       m.n_sites, m.hidden, m.act, cfg, W.W1, W.b1, z1, h1);
  VMCX_LAUNCHED();
  return 0;
}

template <typename TW, typename TI>
static int layer1(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows, const int64_t *cfg,
                  hipStream_t st) {
  return layer1_to<TW, TI>(m, rows, cfg, (TW *)L.z1, (TI *)L.h1, st);
}

// the BF16 forward on the hand-written matrix-core layers: usable when the
// hidden width is a multiple of 16 that fits the LDS staging (the input rows
// of 64 configurations), for N <= 64
static bool mfma_forward_ok(const vmcx_nnbf_model &m) {
  return m.gemm == 1 && m.hidden % 16 == 0 && m.n_sites <= 64 &&
         ng_lds_bytes(m.hidden) <= max_lds_per_block();
}

template <int EPI>
static int launch_ng(int64_t R, int n_rows, int K, const void *W, const void *hin,
                     const float *bias, int act, float *z, void *hout, const int64_t *cfg,
                     int N, int nf, float *A, hipStream_t st) {
  // a batch of few 64-configuration blocks (the proposals of one step)
  // splits W's tiles over blockIdx.y, keeping >= 4 tiles (one per wave) per
  // workgroup, until about 1024 workgroups are in flight
  const int n_tiles = (n_rows + 31) / 32;
  if (EPI == NG_SCAT && K == NG2_K) {
    // 256 configurations per workgroup; W's tiles split over blockIdx.y
    // (at least 4 tiles per slice) until about 512 workgroups are in flight
    const int64_t rb = (R + NG2_CPB - 1) / NG2_CPB;
    int S = 1;
    while (rb * S < 512 && S * 2 * 4 <= n_tiles) S *= 2;
    hipLaunchKernelGGL(nnbf_ng2_kernel, dim3((unsigned)rb, (unsigned)S), dim3(NG2_T), 0, st, R,
                       n_rows, (const __hip_bfloat16 *)W, (const __hip_bfloat16 *)hin, bias, cfg,
                       N, nf, A);
    VMCX_LAUNCHED();
    return 0;
  }
  const int64_t rb = (R + NG_RB - 1) / NG_RB;
  int S = 1;
  while (rb * S < 1024 && S * 2 * 4 <= n_tiles) S *= 2;
  const dim3 g((unsigned)rb, (unsigned)S);
  const size_t lds = ng_lds_bytes(K);
  if (K == 256)
    hipLaunchKernelGGL((nnbf_gemm_mfma_kernel<EPI, 256>), g, dim3(NG_T), lds, st, R, n_rows, K,
                       (const __hip_bfloat16 *)W, (const __hip_bfloat16 *)hin, bias, act, z,
                       (__hip_bfloat16 *)hout, cfg, N, nf, A);
  else
    hipLaunchKernelGGL((nnbf_gemm_mfma_kernel<EPI, 0>), g, dim3(NG_T), lds, st, R, n_rows, K,
                       (const __hip_bfloat16 *)W, (const __hip_bfloat16 *)hin, bias, act, z,
                       (__hip_bfloat16 *)hout, cfg, N, nf, A);
  VMCX_LAUNCHED();
  return 0;
}

template <int NFP>
static void launch_detA(const vmcx_nnbf_model &m, int64_t rows, const int64_t *cfg,
                        const float *A, double *sign, double *logabs, hipStream_t st) {
  constexpr int WPW = 64 / NFP;
  const int64_t per_block = 4 * WPW;
  const int64_t nblk = (rows + per_block - 1) / per_block;
  if (m.det_fp32)
    hipLaunchKernelGGL((nnbf_detA_kernel<NFP, float>), dim3((unsigned)nblk), dim3(256),
                       (size_t)per_block * NFP * 8, st, rows, m.n_sites, m.n_f, cfg, A,
                       (const float *)m.M, m.bf_scale, sign, logabs);
  else
    hipLaunchKernelGGL((nnbf_detA_kernel<NFP, double>), dim3((unsigned)nblk), dim3(256),
                       (size_t)per_block * NFP * 8, st, rows, m.n_sites, m.n_f, cfg, A,
                       (const float *)m.M, m.bf_scale, sign, logabs);
}

// layer 1, layer 2 and the occupied-row output layer on the matrix cores,
// then the determinants: no rocBLAS call and no backflow matrix in HBM
static int forward_rows_mfma(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows,
                             const int64_t *cfg, double *sign, double *logabs, bool h1_ready,
                             hipStream_t st) {
  int rc = h1_ready ? 0 : layer1<float, __hip_bfloat16>(m, L, rows, cfg, st);
  if (rc) return rc;
  const int H = m.hidden, nD = 2 * m.n_sites * m.n_f;
  // (the pre-activations z2 are not kept: only the backward reads them, and
  // it runs its own forward)
  rc = launch_ng<NG_ACT>(rows, H, H, m.W2_bf16, L.h1, m.b2, m.act, nullptr, L.h2, nullptr,
                         m.n_sites, m.n_f, nullptr, st);
  if (rc) return rc;
  float *A = (float *)L.D;   // [rows][nf][nf] fits the [rows][2N nf] buffer
  rc = launch_ng<NG_SCAT>(rows, nD, H, m.Wout_bf16, L.h2, m.bout, m.act, nullptr, nullptr, cfg,
                          m.n_sites, m.n_f, A, st);
  if (rc) return rc;
  switch (nfp_of(m.n_f)) {
    case 8: launch_detA<8>(m, rows, cfg, A, sign, logabs, st); break;
    case 16: launch_detA<16>(m, rows, cfg, A, sign, logabs, st); break;
    case 32: launch_detA<32>(m, rows, cfg, A, sign, logabs, st); break;
    default: launch_detA<64>(m, rows, cfg, A, sign, logabs, st); break;
  }
  VMCX_LAUNCHED();
  return 0;
}

// forward of `rows` configurations (rows <= chunk); ainv optional
template <typename TW, typename TI>
static int forward_rows_t(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows,
                          const int64_t *cfg, double *sign, double *logabs, double *ainv,
                          bool h1_ready, hipStream_t st) {
  if (rows == 0) return 0;
  int rc = h1_ready ? 0 : layer1<TW, TI>(m, L, rows, cfg, st);
  if (rc) return rc;
  rc = dense_layers<TW, TI>(m, L, rows, st);
  if (rc) return rc;
  switch (nfp_of(m.n_f)) {
    case 8: launch_det<8, TW>(m, L, rows, cfg, sign, logabs, ainv, st); break;
    case 16: launch_det<16, TW>(m, L, rows, cfg, sign, logabs, ainv, st); break;
    case 32: launch_det<32, TW>(m, L, rows, cfg, sign, logabs, ainv, st); break;
    default: launch_det<64, TW>(m, L, rows, cfg, sign, logabs, ainv, st); break;
  }
  VMCX_LAUNCHED();
  return 0;
}

// h1_ready: the first layer's output is already in L.h1 (local energy:
// formed incrementally from the walkers'), so the pass starts at layer 2
static int forward_rows(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows,
                        const int64_t *cfg, double *sign, double *logabs, double *ainv,
                        hipStream_t st, bool h1_ready = false) {
  if (m.gemm == 2)
    return forward_rows_t<double, double>(m, L, rows, cfg, sign, logabs, ainv, h1_ready, st);
  if (m.gemm == 1 && ainv == nullptr && rows > 0 && mfma_forward_ok(m) && m.path == 0)
    return forward_rows_mfma(m, L, rows, cfg, sign, logabs, h1_ready, st);
  if (m.gemm == 1)
    return forward_rows_t<float, __hip_bfloat16>(m, L, rows, cfg, sign, logabs, ainv, h1_ready,
                                                 st);
  return forward_rows_t<float, float>(m, L, rows, cfg, sign, logabs, ainv, h1_ready, st);
}

// walkers' first-layer pre-activations into L.z1w (in chunks through L.h1)
// and W1 transposed into L.w1t: the inputs of the incremental first layer
template <typename TW, typename TI>
static int conn_layer1_setup(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t B, int64_t chunk,
                             const int64_t *cfg, hipStream_t st) {
  const int H = m.hidden, n_orb = 2 * m.n_sites;
  for (int64_t r0 = 0; r0 < B; r0 += chunk) {
    const int64_t rows = (B - r0 < chunk) ? B - r0 : chunk;
    const int rc = layer1_to<TW, TI>(m, rows, cfg + r0 * m.n_sites, (TW *)L.z1w + r0 * H,
                                     (TI *)L.h1, st);
    if (rc) return rc;
  }
  const NnW<TW> W = weights_of<TW>(m);
  const int64_t n = (int64_t)H * n_orb;
  hipLaunchKernelGGL((nnbf_w1t_kernel<TW>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     H, n_orb, W.W1, (TW *)L.w1t);
  VMCX_LAUNCHED();
  return 0;
}

template <typename TW, typename TI>
static int conn_layer1(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t r0, int64_t rows,
                       hipStream_t st) {
  const int64_t n = rows * m.hidden;
  const unsigned g = (unsigned)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384);
  hipLaunchKernelGGL((nnbf_h1_conn_kernel<TW, TI>), dim3(g), dim3(256), 0, st, rows, m.hidden,
                     (const TW *)L.w1t, (const TW *)L.z1w, L.cparent + r0, L.cmove + r0, m.act,
                     (TI *)L.h1);
  VMCX_LAUNCHED();
  return 0;
}

template <typename TW, typename TI>
static int moved_layer1(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows,
                        const int64_t *cfg, const int64_t *pcfg, const void *z1, void *z1p,
                        hipStream_t st) {
  hipLaunchKernelGGL((nnbf_h1_moved_kernel<TW, TI>), dim3((unsigned)((rows + 3) / 4)), dim3(256),
                     0, st, rows, m.n_sites, m.hidden, (const TW *)L.w1t, (const TW *)z1, cfg,
                     pcfg, m.act, (TW *)z1p, (TI *)L.h1);
  VMCX_LAUNCHED();
  return 0;
}

int nnbf_forward_moved(const vmcx_nnbf_model *m, int64_t B, const int64_t *cfg,
                       const int64_t *pcfg, const void *z1, void *ws, int64_t ws_bytes,
                       int64_t chunk, double *sign, double *logabs, void *z1p, hipStream_t st) {
  if (check_model(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && chunk >= 1, "bad sizes");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(cfg && pcfg && z1 && z1p && sign && logabs && ws, "null pointer");
  const NnbfWs L = nnbf_layout(*m, B, chunk, 0, (char *)ws);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes, "workspace too small");
  const int H = m->hidden, n_orb = 2 * m->n_sites, N = m->n_sites;
  const size_t tw = m->gemm == 2 ? 8 : 4;
  {
    const int64_t n = (int64_t)H * n_orb;
    if (m->gemm == 2)
      hipLaunchKernelGGL((nnbf_w1t_kernel<double>), dim3((unsigned)((n + 255) / 256)), dim3(256),
                         0, st, H, n_orb, weights_of<double>(*m).W1, (double *)L.w1t);
    else
      hipLaunchKernelGGL((nnbf_w1t_kernel<float>), dim3((unsigned)((n + 255) / 256)), dim3(256),
                         0, st, H, n_orb, weights_of<float>(*m).W1, (float *)L.w1t);
    VMCX_LAUNCHED();
  }
  for (int64_t r0 = 0; r0 < B; r0 += chunk) {
    const int64_t rows = (B - r0 < chunk) ? B - r0 : chunk;
    const char *zi = (const char *)z1 + (size_t)r0 * H * tw;
    char *zo = (char *)z1p + (size_t)r0 * H * tw;
    int rc;
    if (m->gemm == 2)
      rc = moved_layer1<double, double>(*m, L, rows, cfg + r0 * N, pcfg + r0 * N, zi, zo, st);
    else if (m->gemm == 1)
      rc = moved_layer1<float, __hip_bfloat16>(*m, L, rows, cfg + r0 * N, pcfg + r0 * N, zi, zo,
                                               st);
    else
      rc = moved_layer1<float, float>(*m, L, rows, cfg + r0 * N, pcfg + r0 * N, zi, zo, st);
    if (rc) return rc;
    rc = forward_rows(*m, L, rows, pcfg + r0 * N, sign + r0, logabs + r0, nullptr, st, true);
    if (rc) return rc;
  }
  return 0;
}

int masked_rows(int64_t B, int64_t row_bytes, const uint8_t *mask, const void *src, void *dst,
                hipStream_t st) {
  VMCX_CHECK_ARG(B >= 0 && row_bytes >= 0 && row_bytes % 4 == 0,
                 "row_bytes must be a multiple of 4");
  if (B == 0 || row_bytes == 0) return 0;
  VMCX_CHECK_ARG(mask && src && dst, "null pointer");
  const int64_t n = B * (row_bytes / 4);
  const unsigned g = (unsigned)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384);
  hipLaunchKernelGGL(masked_rows_kernel, dim3(g), dim3(256), 0, st, B, row_bytes / 4, mask,
                     (const uint32_t *)src, (uint32_t *)dst);
  VMCX_LAUNCHED();
  return 0;
}

int nnbf_forward(const vmcx_nnbf_model *m, int64_t B, const int64_t *cfg, void *ws,
                 int64_t ws_bytes, int64_t chunk, double *sign, double *logabs, void *z1_out,
                 hipStream_t st) {
  if (check_model(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && chunk >= 1, "bad sizes");
  VMCX_CHECK_ARG(B == 0 || (cfg && sign && logabs && ws), "null pointer");
  const NnbfWs L = nnbf_layout(*m, B, chunk, 0, (char *)ws);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes, "workspace too small");
  const size_t tw = m->gemm == 2 ? 8 : 4;
  for (int64_t r0 = 0; r0 < B; r0 += chunk) {
    const int64_t rows = (B - r0 < chunk) ? B - r0 : chunk;
    const int rc = forward_rows(*m, L, rows, cfg + r0 * m->n_sites, sign + r0, logabs + r0,
                                nullptr, st);
    if (rc) return rc;
    if (z1_out)
      VMCX_HIP(hipMemcpyAsync((char *)z1_out + (size_t)r0 * m->hidden * tw, L.z1,
                              (size_t)rows * m->hidden * tw, hipMemcpyDeviceToDevice, st));
  }
  return 0;
}

template <int NFP, typename TW>
static void launch_accept(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t B, int i, int j,
                          int64_t *cfg, TW *z1, double *sign, double *logabs,
                          int64_t *n_accept, hipStream_t st) {
  constexpr int WPW = 64 / NFP;
  const int64_t per_block = 4 * WPW;
  const int64_t nblk = (B + per_block - 1) / per_block;
  const NnW<TW> W = weights_of<TW>(m);
  hipLaunchKernelGGL((nnbf_accept_kernel<NFP, TW>), dim3((unsigned)nblk), dim3(256),
                     (size_t)per_block * NFP * 9, st, B, m.n_sites, m.n_f, m.hidden, i, j, cfg,
                     L.prop, L.moves, L.uacc, (const TW *)L.D, W.M, W.bout, m.bf_scale, W.W1,
                     z1, sign, logabs, n_accept);
}

template <typename TW, typename TI>
static int sweep_edge(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t B, int i, int j,
                      int64_t *cfg, TW *z1, double *sign, double *logabs, int64_t *n_accept,
                      hipStream_t st) {
  const int64_t nh = B * m.hidden;
  const unsigned gh = (unsigned)((nh + 255) / 256 < 16384 ? (nh + 255) / 256 : 16384);
  const NnW<TW> W = weights_of<TW>(m);
  hipLaunchKernelGGL((nnbf_h1_proposed_kernel<TW, TI>), dim3(gh), dim3(256), 0, st, B,
                     m.n_sites, m.hidden, m.act, W.W1, z1, L.prop, L.moves, (TI *)L.h1);
  VMCX_LAUNCHED();
  const int rc = dense_layers<TW, TI>(m, L, B, st);
  if (rc) return rc;
  switch (nfp_of(m.n_f)) {
    case 8: launch_accept<8, TW>(m, L, B, i, j, cfg, z1, sign, logabs, n_accept, st); break;
    case 16: launch_accept<16, TW>(m, L, B, i, j, cfg, z1, sign, logabs, n_accept, st); break;
    case 32: launch_accept<32, TW>(m, L, B, i, j, cfg, z1, sign, logabs, n_accept, st); break;
    default: launch_accept<64, TW>(m, L, B, i, j, cfg, z1, sign, logabs, n_accept, st); break;
  }
  VMCX_LAUNCHED();
  return 0;
}

int nnbf_sweep(const vmcx_nnbf_model *m, int64_t B, int n_edges, const int32_t *edges_host,
               double hopping_rate, int n_sweeps, uint64_t seed, int64_t sweep_offset,
               const double *u_choice, const uint8_t *bits, const double *u_acc,
               int64_t *cfg, void *z1, double *sign, double *logabs, int64_t *n_accept,
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
  for (int sw = 0; sw < n_sweeps; ++sw) {
    for (int e = 0; e < n_edges; ++e) {
      const int i = edges_host[2 * e], j = edges_host[2 * e + 1];
      const size_t di = ((size_t)sw * n_edges + e) * B;
      hipLaunchKernelGGL(nnbf_propose_kernel, dim3(gw), dim3(256), 0, st, B, N, i, j, e,
                         sweep_offset + sw, seed, 1.0 - hopping_rate,
                         u_choice ? u_choice + di : nullptr, bits ? bits + di : nullptr,
                         u_acc ? u_acc + di : nullptr, cfg, L.prop, L.moves, L.uacc);
      VMCX_LAUNCHED();
      int rc;
      if (m->gemm == 2)
        rc = sweep_edge<double, double>(*m, L, B, i, j, cfg, (double *)z1, sign, logabs,
                                        n_accept, st);
      else if (m->gemm == 1)
        rc = sweep_edge<float, __hip_bfloat16>(*m, L, B, i, j, cfg, (float *)z1, sign, logabs,
                                               n_accept, st);
      else
        rc = sweep_edge<float, float>(*m, L, B, i, j, cfg, (float *)z1, sign, logabs, n_accept,
                                      st);
      if (rc) return rc;
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
                                    st, L.cparent, L.cmove);
    if (rc) return rc;
  }
  // the first layer of every connected configuration from its walker's
  // (one hop = one column pair of W1), the rest of the pass per chunk
  {
    int rc;
    if (m->gemm == 2) rc = conn_layer1_setup<double, double>(*m, L, B, chunk, cfg, st);
    else if (m->gemm == 1) rc = conn_layer1_setup<float, __hip_bfloat16>(*m, L, B, chunk, cfg, st);
    else rc = conn_layer1_setup<float, float>(*m, L, B, chunk, cfg, st);
    if (rc) return rc;
  }
  int64_t T = 0;
  VMCX_HIP(hipMemcpyAsync(&T, L.total, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  VMCX_HIP(hipStreamSynchronize(st));
  for (int64_t r0 = 0; r0 < T; r0 += chunk) {
    const int64_t rows = (T - r0 < chunk) ? T - r0 : chunk;
    int rc;
    if (m->gemm == 2) rc = conn_layer1<double, double>(*m, L, r0, rows, st);
    else if (m->gemm == 1) rc = conn_layer1<float, __hip_bfloat16>(*m, L, r0, rows, st);
    else rc = conn_layer1<float, float>(*m, L, r0, rows, st);
    if (rc) return rc;
    rc = forward_rows(*m, L, rows, L.etas + r0 * N, L.csign + r0, L.clog + r0, nullptr, st,
                      true);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(eloc_segment_kernel, dim3(gw), dim3(256), 0, st, B, N, U, cfg, L.offsets,
                     L.counts, L.coeffs, L.csign, L.clog, sign, logabs, eloc);
  VMCX_LAUNCHED();
  return 0;
}

// rows of O (or of its factors) for `rows` configurations whose forward
// pass left z1, z2 and A^-1 in the workspace
template <typename TO, bool FACT, typename TW>
static int backward_rows_t(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows,
                           const int64_t *cfg, TO *O, int64_t ld, hipStream_t st) {
  const int N = m.n_sites, H = m.hidden, nf = m.n_f, nD = 2 * N * nf;
  rocblas_handle h = blas_handle_for(st);
  VMCX_CHECK_ARG(h != nullptr, "rocblas handle");
  const NnW<TW> W = weights_of<TW>(m);
  hipLaunchKernelGGL((nnbf_grad_head_kernel<TO, TW>), dim3((unsigned)rows), dim3(256), 0, st,
                     rows, N, nf, nfp_of(nf), cfg, L.ainv, m.bf_scale, (TW *)L.D, O, ld);
  VMCX_LAUNCHED();
  const int ri = (int)rows;
  const int gm = m.gemm == 2 ? 2 : 0;   // FP64 or FP32 products (the backward is FP32 in BF16 mode)
  // column-major views: dh2 (H x rows) = Wout^T (Wout row-major [nD][H]) * (s gM) (nD x rows)
  int rc = nn_gemm(h, rocblas_operation_none, H, ri, nD, W.Wout, H, L.D, nD, L.g2, H, gm);
  if (rc) return rc;
  const int64_t nh = rows * H;
  const unsigned gh = (unsigned)((nh + 255) / 256 < 8192 ? (nh + 255) / 256 : 8192);
  hipLaunchKernelGGL(nnbf_dact_kernel<TW>, dim3(gh), dim3(256), 0, st, nh, m.act,
                     (const TW *)L.z2, (TW *)L.g2);
  VMCX_LAUNCHED();
  // dh1 (H x rows) = W2^T-as-stored (W2 row-major [u][i]: column-major (i, u)) * dz2
  rc = nn_gemm(h, rocblas_operation_none, H, ri, H, W.W2, H, L.g2, H, L.g1, H, gm);
  if (rc) return rc;
  hipLaunchKernelGGL(nnbf_dact_kernel<TW>, dim3(gh), dim3(256), 0, st, nh, m.act,
                     (const TW *)L.z1, (TW *)L.g1);
  VMCX_LAUNCHED();
  const size_t lds = (size_t)(4 * H + 2 * N) * sizeof(TW);
  hipLaunchKernelGGL((nnbf_grad_tail_kernel<TO, FACT, TW>), dim3((unsigned)rows), dim3(256), lds,
                     st, rows, N, nf, H, m.act, cfg, (const TW *)L.z1, (const TW *)L.z2,
                     (const TW *)L.g1, (const TW *)L.g2, (const TW *)L.D, O, ld);
  VMCX_LAUNCHED();
  return 0;
}

template <typename TO, bool FACT>
static int backward_rows(const vmcx_nnbf_model &m, const NnbfWs &L, int64_t rows,
                         const int64_t *cfg, TO *O, int64_t ld, hipStream_t st) {
  if (m.gemm == 2) return backward_rows_t<TO, FACT, double>(m, L, rows, cfg, O, ld, st);
  return backward_rows_t<TO, FACT, float>(m, L, rows, cfg, O, ld, st);
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
  const size_t tw = m->gemm == 2 ? 8 : 4;
  VMCX_CHECK_ARG((4 * H + 2 * N) * tw <= 65536, "backward LDS exceeds 64 KB");
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
  const size_t tw = m->gemm == 2 ? 8 : 4;
  VMCX_CHECK_ARG((4 * H + 2 * N) * tw <= 65536, "backward LDS exceeds 64 KB");
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
