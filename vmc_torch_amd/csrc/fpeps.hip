// Z2-symmetric fermionic PEPS (fPEPS) and NN-fTNS walkers on gfx950.
//
// Amplitude of a configuration x (one workgroup of 256 threads per
// configuration):
//   psi(x) = (-1)^{C(P,2)} contract_s S_s,
//   S_s[l,u,d,r] = T_s[x_s,l,u,d,r] (-1)^{q(d) (q(r) + P_right(s))},
// P = number of singly occupied sites, P_right(s) = those right of s in its
// row, q(k) = parity of virtual index k (k >= D/2; dimension-1 boundary legs
// even).  This is the Grassmann tensor network with modes in the Hamiltonian's
// Jordan-Wigner order (sites row-major, down before up) and leg order
// (l, u, d, r); the local factor is what the crossings of odd bonds reduce to.
//
// Contraction (exact for Lx <= 4 and bond dimension <= 4, the top and bottom
// halves being at most two rows): column by column, the top half (rows
// 0 .. Lx/2-1) and the bottom half (rows Lx/2 .. Lx-1) of column y are
// combined into Top_y[a][v][a'] and Bot_y[b][v][b'] (a, b: left bonds of the
// half, v: the vertical bond between the halves), and the environment
//   env_{y+1}[a'][b'] = sum_{a,b,v} env_y[a][b] Top_y[a][v][a'] Bot_y[b][v][b']
// is renormalised after every column (log scale carried).  The log-derivative
// kernel keeps every env_y, sweeps back from the right with the right
// environment and differentiates Top/Bot, then the rows inside each half.
//
// NN-fTNS: theta(x) = theta + eta * MLP(n(x)) (additive, the reference's
// form) or theta(x) = theta * (1 + eta * MLP(n(x))) (multiplicative), n = [up |
// down] occupations;
// first layer as a gather of W1 columns, the dense layers as rocBLAS DGEMMs,
// every tensor entry then read per configuration from theta(x).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cmath>

#include "kernels.h"

namespace vmcx {

namespace {

constexpr int PMAX_LX = 4, PMAX_LY = 8, PMAX_D = 4, PMAX_N = PMAX_LX * PMAX_LY;
constexpr int PMAX_NS = 64;         // sites of a boundary-MPS lattice (the MLP kernels)
constexpr int PB = 16;              // max bond of a half (D^2)
constexpr int PT = 256;             // threads per configuration
constexpr int SLICE = PMAX_D * PMAX_D * PMAX_D * PMAX_D;

struct PepsGeo {
  int Lx, Ly, D, N, tr;  // tr = rows in the top half
};

__device__ __forceinline__ int qpar(int k, int d) { return d == 1 ? 0 : (k >= (d >> 1)); }

__device__ __forceinline__ void site_dims(const PepsGeo &g, int x, int y, int &dl, int &du,
                                          int &dd, int &dr) {
  dl = (y == 0) ? 1 : g.D;
  du = (x == 0) ? 1 : g.D;
  dd = (x == g.Lx - 1) ? 1 : g.D;
  dr = (y == g.Ly - 1) ? 1 : g.D;
}

// per-configuration data shared by the workgroup
struct CfgSh {
  int p[PMAX_N];      // local states
  int pr[PMAX_N];     // P_right parity
  int doff[PMAX_N];   // dense offset of each site tensor
  double gsign;       // (-1)^{C(P,2)}
};

__device__ void load_cfg(const PepsGeo &g, const int64_t *c, CfgSh &cs) {
  const int tid = threadIdx.x;
  if (tid < g.N) cs.p[tid] = (int)c[tid];
  __syncthreads();
  if (tid < g.N) {
    const int x = tid / g.Ly, y = tid - x * g.Ly;
    int pr = 0;
    for (int yy = y + 1; yy < g.Ly; ++yy) {
      const int v = cs.p[x * g.Ly + yy];
      pr += (v == 1 || v == 2);
    }
    cs.pr[tid] = pr & 1;
  }
  if (tid == 0) {
    int P = 0, off = 0;
    for (int s = 0; s < g.N; ++s) {
      const int v = cs.p[s];
      P += (v == 1 || v == 2);
      const int x = s / g.Ly, y = s - x * g.Ly;
      int dl, du, dd, dr;
      site_dims(g, x, y, dl, du, dd, dr);
      cs.doff[s] = off;
      off += 4 * dl * du * dd * dr;
    }
    cs.gsign = ((P * (P - 1) / 2) & 1) ? -1.0 : 1.0;
  }
  __syncthreads();
}

// signed slice of site (x, y) into dst[l][u][d][r] (dense dims)
__device__ void load_slice(const PepsGeo &g, const CfgSh &cs, int x, int y,
                           const int32_t *__restrict__ dmap, const double *__restrict__ th,
                           double *dst) {
  int dl, du, dd, dr;
  site_dims(g, x, y, dl, du, dd, dr);
  const int s = x * g.Ly + y;
  const int n = dl * du * dd * dr;
  const int base = cs.doff[s] + cs.p[s] * n;
  const int pr = cs.pr[s];
  for (int idx = threadIdx.x; idx < n; idx += PT) {
    const int r = idx % dr, d = (idx / dr) % dd;
    const int k = dmap[base + idx];
    double v = k >= 0 ? th[k] : 0.0;
    if ((qpar(d, dd) * (qpar(r, dr) + pr)) & 1) v = -v;
    dst[idx] = v;
  }
}

struct ColSh {
  double sl[PMAX_LX][SLICE];        // signed slices of the column
  double top[PB * PMAX_D * PB];     // Top[a][v][a']
  double bot[PB * PMAX_D * PB];     // Bot[b][v][b']
  double X[PB * PMAX_D * PB];
  int da, da2, db, db2, dv;
};

// slices, Top and Bot of column y
__device__ void build_column(const PepsGeo &g, const CfgSh &cs, int y,
                             const int32_t *dmap, const double *th, ColSh &c) {
  for (int x = 0; x < g.Lx; ++x) load_slice(g, cs, x, y, dmap, th, c.sl[x]);
  const int D = g.D;
  const int dlc = (y == 0) ? 1 : D, drc = (y == g.Ly - 1) ? 1 : D;
  const int br = g.Lx - g.tr;
  const int da = g.tr == 2 ? dlc * dlc : dlc, da2 = g.tr == 2 ? drc * drc : drc;
  const int db = br == 2 ? dlc * dlc : dlc, db2 = br == 2 ? drc * drc : drc;
  const int dv = D;
  if (threadIdx.x == 0) {
    c.da = da; c.da2 = da2; c.db = db; c.db2 = db2; c.dv = dv;
  }
  __syncthreads();
  // top
  const int ntop = da * dv * da2;
  for (int idx = threadIdx.x; idx < ntop; idx += PT) {
    const int a2 = idx % da2, v = (idx / da2) % dv, a = idx / (da2 * dv);
    double acc;
    if (g.tr == 1) {
      // row 0: [l][u=0][d=v][r]
      acc = c.sl[0][(a * dv + v) * drc + a2];
    } else {
      const int l0 = a / dlc, l1 = a % dlc, r0 = a2 / drc, r1 = a2 % drc;
      // row 0 [l0][0][m][r0], row 1 [l1][m][v][r1]
      acc = 0.0;
      for (int m = 0; m < D; ++m)
        acc += c.sl[0][(l0 * D + m) * drc + r0] * c.sl[1][((l1 * D + m) * dv + v) * drc + r1];
    }
    c.top[idx] = acc;
  }
  // bottom
  const int nbot = db * dv * db2;
  const int xb = g.tr;  // first bottom row
  for (int idx = threadIdx.x; idx < nbot; idx += PT) {
    const int b2 = idx % db2, v = (idx / db2) % dv, b = idx / (db2 * dv);
    double acc;
    if (br == 1) {
      // last row [l][u=v][d=0][r]
      acc = c.sl[xb][(b * dv + v) * drc + b2];
    } else {
      const int l2 = b / dlc, l3 = b % dlc, r2 = b2 / drc, r3 = b2 % drc;
      // row xb [l2][v][m][r2], row xb+1 [l3][m][0][r3]
      acc = 0.0;
      for (int m = 0; m < D; ++m)
        acc += c.sl[xb][((l2 * dv + v) * D + m) * drc + r2] *
               c.sl[xb + 1][(l3 * D + m) * drc + r3];
    }
    c.bot[idx] = acc;
  }
  __syncthreads();
}

// env_out[a'][b'] = sum env[a][b] Top[a][v][a'] Bot[b][v][b'] (strides PB)
__device__ void absorb_left(ColSh &c, const double *env, double *env_out) {
  const int da = c.da, da2 = c.da2, db = c.db, db2 = c.db2, dv = c.dv;
  const int nx = db * dv * da2;
  for (int idx = threadIdx.x; idx < nx; idx += PT) {
    const int a2 = idx % da2, v = (idx / da2) % dv, b = idx / (da2 * dv);
    double acc = 0.0;
    for (int a = 0; a < da; ++a) acc += env[a * PB + b] * c.top[(a * dv + v) * da2 + a2];
    c.X[idx] = acc;  // X[b][v][a']
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < da2 * db2; idx += PT) {
    const int b2 = idx % db2, a2 = idx / db2;
    double acc = 0.0;
    for (int b = 0; b < db; ++b)
      for (int v = 0; v < dv; ++v)
        acc += c.X[(b * dv + v) * da2 + a2] * c.bot[(b * dv + v) * db2 + b2];
    env_out[a2 * PB + b2] = acc;
  }
  __syncthreads();
}

// R_out[a][b] = sum Top[a][v][a'] Bot[b][v][b'] R[a'][b']
__device__ void absorb_right(ColSh &c, const double *R, double *R_out) {
  const int da = c.da, da2 = c.da2, db = c.db, db2 = c.db2, dv = c.dv;
  // Z[a][v][b'] = sum_a' Top[a][v][a'] R[a'][b']
  for (int idx = threadIdx.x; idx < da * dv * db2; idx += PT) {
    const int b2 = idx % db2, v = (idx / db2) % dv, a = idx / (db2 * dv);
    double acc = 0.0;
    for (int a2 = 0; a2 < da2; ++a2) acc += c.top[(a * dv + v) * da2 + a2] * R[a2 * PB + b2];
    c.X[idx] = acc;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < da * db; idx += PT) {
    const int b = idx % db, a = idx / db;
    double acc = 0.0;
    for (int v = 0; v < dv; ++v)
      for (int b2 = 0; b2 < db2; ++b2)
        acc += c.X[(a * dv + v) * db2 + b2] * c.bot[(b * dv + v) * db2 + b2];
    R_out[a * PB + b] = acc;
  }
  __syncthreads();
}

// divide env (da x db, stride PB) by its largest magnitude; returns log of it
__device__ double renorm(double *env, int da, int db, double *red) {
  double m = 0.0;
  for (int idx = threadIdx.x; idx < da * db; idx += PT) {
    const int a = idx / db, b = idx % db;
    m = fmax(m, fabs(env[a * PB + b]));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  double mm = 0.0;
  for (int k = 0; k < PT / 64; ++k) mm = fmax(mm, red[k]);
  __syncthreads();
  if (mm > 0.0) {
    const double inv = 1.0 / mm;
    for (int idx = threadIdx.x; idx < da * db; idx += PT) {
      const int a = idx / db, b = idx % db;
      env[a * PB + b] *= inv;
    }
  }
  __syncthreads();
  return mm > 0.0 ? log(mm) : 0.0;
}

struct PepsArgs {
  PepsGeo g;
  const int32_t *dmap;
  const double *theta;
  int64_t theta_stride;   // per-configuration theta rows (0: shared)
  int64_t n_params;
};

}  // namespace

// ------------------------------------------------------------ amplitudes
// skip (may be NULL): configuration rows with !(skip[row] & 1) are left alone
__global__ void __launch_bounds__(PT)
    peps_amp_kernel(PepsArgs A, int64_t rows, const int64_t *__restrict__ cfg,
                    const int32_t *__restrict__ flag, double *sign, double *logabs) {
  const int64_t w = blockIdx.x;
  if (w >= rows) return;
  if (flag && !(flag[w] & 1)) return;
  __shared__ CfgSh cs;
  __shared__ ColSh col;
  __shared__ double env[2][PB * PB];
  __shared__ double red[PT / 64];
  const PepsGeo &g = A.g;
  load_cfg(g, cfg + w * g.N, cs);
  const double *th = A.theta + w * A.theta_stride;
  if (threadIdx.x == 0) env[0][0] = 1.0;
  __syncthreads();
  double lsc = 0.0;
  int cur = 0;
  for (int y = 0; y < g.Ly; ++y) {
    build_column(g, cs, y, A.dmap, th, col);
    absorb_left(col, env[cur], env[cur ^ 1]);
    cur ^= 1;
    lsc += renorm(env[cur], col.da2, col.db2, red);
  }
  if (threadIdx.x == 0) {
    const double v = env[cur][0] * cs.gsign;
    const double lo = log(1e-45);
    if (v == 0.0 || !std::isfinite(v)) {
      sign[w] = 0.0;
      logabs[w] = lo;
    } else {
      sign[w] = v > 0.0 ? 1.0 : -1.0;
      logabs[w] = fmax(log(fabs(v)) + lsc, lo);
    }
  }
}

// ------------------------------------------------------------ FP32 amplitudes
// The same contraction in FP32 on the matrix cores, one wave per
// configuration (four per workgroup).  Top, Bot and the environment are
// zero-padded to 16 x 4 x 16 and 16 x 16 so that both products of a column
// are whole v_mfma_f32_16x16x4f32 tiles:
//   X[b][v][a'] = sum_a env[a][b] Top[a][v][a']     (A = env^T, B = Top[., v, .]: 4 tiles x 4 k-steps)
//   env'[a'][b'] = sum_{b,v} X[b][v][a'] Bot[b][v][b'] (A = X^T, B = Bot: 16 k-steps)
// (operands, lane l: A[l % 16][l / 16], B[l / 16][l % 16]; result register
// r of lane l: D[4 (l / 16) + r][l % 16]).  The slices are read from the FP64
// parameters and rounded once; the environment is renormalised after every
// column (its log scale carried in FP64), as in peps_amp_kernel.
constexpr int P32_W = 4;
typedef float p32f4 __attribute__((ext_vector_type(4)));

// the wave's LDS writes visible to its own later reads
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct P32Sh {
  float sl[PMAX_LX][SLICE];          // signed slices of the column; then X[b][v][a']
  float top[PB * PMAX_D * PB];       // Top[a][v][a'], padded
  float bot[PB * PMAX_D * PB];       // Bot[b][v][b'], padded
  float env[PB * PB];                // env[a][b], padded
  int p[PMAX_N], pr[PMAX_N], doff[PMAX_N];
};

// Signed FP32 slices of a shared parameter vector (plain fPEPS), for every
// site s, local state p and parity pr of P_right(s): tbl[((s 4 + p) 2 + pr)
// 256 + idx] (zero past the slice's dl du dd dr entries)
__global__ void __launch_bounds__(256)
    peps_slice_table_kernel(PepsArgs A, float *tbl) {
  const PepsGeo &g = A.g;
  const int64_t n_all = (int64_t)g.N * 8 * 256;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_all;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int idx = (int)(e & 255), q = (int)(e >> 8);
    const int prs = q & 1, pst = (q >> 1) & 3, s = q >> 3;
    const int x = s / g.Ly, y = s - (s / g.Ly) * g.Ly;
    int off = 0;
    for (int s2 = 0; s2 < s; ++s2) {
      int a, b, c, d;
      site_dims(g, s2 / g.Ly, s2 - (s2 / g.Ly) * g.Ly, a, b, c, d);
      off += 4 * a * b * c * d;
    }
    int dl, du, dd, dr;
    site_dims(g, x, y, dl, du, dd, dr);
    const int n = dl * du * dd * dr;
    float v = 0.0f;
    if (idx < n) {
      const int r = idx % dr, d = (idx / dr) % dd;
      const int k = A.dmap[off + pst * n + idx];
      v = k >= 0 ? (float)A.theta[k] : 0.0f;
      if ((qpar(d, dd) * (qpar(r, dr) + prs)) & 1) v = -v;
    }
    tbl[e] = v;
  }
}

// DD: the bond dimension (1 .. 4, compile time); tbl: the slice table of a
// shared parameter vector, or NULL (per-configuration parameters: slices
// read through the dense map)
template <int DD>
__global__ void __launch_bounds__(64 * P32_W)
    peps_amp32_kernel(PepsArgs A, int64_t rows, const int64_t *__restrict__ cfg,
                      const int32_t *__restrict__ flag, const float *__restrict__ tbl,
                      double *sign, double *logabs) {
  __shared__ __align__(16) P32Sh shm[P32_W];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t w = (int64_t)blockIdx.x * P32_W + wv;
  if (w >= rows) return;
  if (flag && !(flag[w] & 1)) return;
  P32Sh &S = shm[wv];
  const PepsGeo &g = A.g;
  const int N = g.N, Ly = g.Ly;
  constexpr int D = DD;
  const double *th = A.theta + w * A.theta_stride;
  // the configuration: local states, P_right parities (singly occupied
  // sites right of s in its row), dense offsets, the global sign
  const int vl = lane < N ? (int)cfg[w * N + lane] : 0;
  const uint64_t single = __ballot(lane < N && (vl == 1 || vl == 2));
  const int P = __popcll(single);
  if (lane < N) {
    const int x = lane / Ly;
    const int cnt = x * Ly + Ly - lane - 1;          // sites right of s in row x
    const uint64_t m = (lane + 1 < 64 ? single >> (lane + 1) : 0ull) &
                       (cnt >= 64 ? ~0ull : ((1ull << cnt) - 1ull));
    S.p[lane] = vl;
    S.pr[lane] = __popcll(m) & 1;
    int off = 0;
    for (int s2 = 0; s2 < lane; ++s2) {
      int dl, du, dd, dr;
      site_dims(g, s2 / Ly, s2 - (s2 / Ly) * Ly, dl, du, dd, dr);
      off += 4 * dl * du * dd * dr;
    }
    S.doff[lane] = off;
  }
  const double gsign = ((P * (P - 1) / 2) & 1) ? -1.0 : 1.0;
  for (int i = lane; i < PB * PB; i += 64) S.env[i] = i == 0 ? 1.0f : 0.0f;
  wave_lds_fence();
  const int tr = g.tr, br = g.Lx - tr, dv = D;
  double lsc = 0.0;
  for (int y = 0; y < Ly; ++y) {
    // signed slices of the column
    if (tbl) {
      // one 16-byte piece of each site's 256-entry table row per lane
      for (int x = 0; x < g.Lx; ++x) {
        const int s = x * Ly + y;
        const float4 v = reinterpret_cast<const float4 *>(
            tbl + ((int64_t)(s * 4 + S.p[s]) * 2 + S.pr[s]) * 256)[lane];
        reinterpret_cast<float4 *>(S.sl[x])[lane] = v;
      }
    } else {
      for (int x = 0; x < g.Lx; ++x) {
        int dl, du, dd, dr;
        site_dims(g, x, y, dl, du, dd, dr);
        const int s = x * Ly + y, n = dl * du * dd * dr;
        const int base = S.doff[s] + S.p[s] * n, prs = S.pr[s];
        for (int idx = lane; idx < n; idx += 64) {
          const int r = idx % dr, d = (idx / dr) % dd;
          const int k = A.dmap[base + idx];
          float v = k >= 0 ? (float)th[k] : 0.0f;
          if ((qpar(d, dd) * (qpar(r, dr) + prs)) & 1) v = -v;
          S.sl[x][idx] = v;
        }
      }
    }
    wave_lds_fence();
    const int dlc = (y == 0) ? 1 : D, drc = (y == Ly - 1) ? 1 : D;
    const int da = tr == 2 ? dlc * dlc : dlc, da2 = tr == 2 ? drc * drc : drc;
    const int db = br == 2 ? dlc * dlc : dlc, db2 = br == 2 ? drc * drc : drc;
    // Top[a][v][a'] and Bot[b][v][b'], zero outside (da or db) x dv x (da2 or db2)
    for (int idx = lane; idx < PB * PMAX_D * PB; idx += 64) {
      const int a = idx >> 6, v = (idx >> 4) & 3, a2 = idx & 15;
      float t = 0.0f, bb = 0.0f;
      if (v < dv) {
        if (a < da && a2 < da2) {
          if (tr == 1) {
            t = S.sl[0][(a * dv + v) * drc + a2];
          } else {
            const int l0 = a / dlc, l1 = a % dlc, r0 = a2 / drc, r1 = a2 % drc;
            for (int m = 0; m < D; ++m)
              t = fmaf(S.sl[0][(l0 * D + m) * drc + r0],
                       S.sl[1][((l1 * D + m) * dv + v) * drc + r1], t);
          }
        }
        if (a < db && a2 < db2) {
          const int xb = tr;
          if (br == 1) {
            bb = S.sl[xb][(a * dv + v) * drc + a2];
          } else {
            const int l2 = a / dlc, l3 = a % dlc, r2 = a2 / drc, r3 = a2 % drc;
            for (int m = 0; m < D; ++m)
              bb = fmaf(S.sl[xb][((l2 * dv + v) * D + m) * drc + r2],
                        S.sl[xb + 1][(l3 * D + m) * drc + r3], bb);
          }
        }
      }
      S.top[idx] = t;
      S.bot[idx] = bb;
    }
    wave_lds_fence();
    // X[b][v][a'] = sum_a env[a][b] Top[a][v][a'] (into the slice buffer)
    float *X = &S.sl[0][0];
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      p32f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const float av = S.env[(4 * ks + lk) * PB + li];
        const float bv = S.top[(4 * ks + lk) * 64 + nb * 16 + li];
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) X[((4 * lk + r) * 4 + nb) * 16 + li] = acc[r];
    }
    wave_lds_fence();
    // env'[a'][b'] = sum_{b,v} X[b][v][a'] Bot[b][v][b']
    p32f4 e = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
      const float av = X[(4 * ks + lk) * 16 + li];
      const float bv = S.bot[(4 * ks + lk) * 16 + li];
      e = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, e, 0, 0, 0);
    }
    // renormalise by the largest magnitude
    float mx = fmaxf(fmaxf(fabsf(e[0]), fabsf(e[1])), fmaxf(fabsf(e[2]), fabsf(e[3])));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    const float inv = mx > 0.0f ? 1.0f / mx : 1.0f;
    if (mx > 0.0f) lsc += log((double)mx);
#pragma unroll
    for (int r = 0; r < 4; ++r) S.env[(4 * lk + r) * PB + li] = e[r] * inv;
    wave_lds_fence();
  }
  if (lane == 0) {
    const double v = (double)S.env[0] * gsign;
    const double lo = log(1e-45);
    if (v == 0.0 || !std::isfinite(v)) {
      sign[w] = 0.0;
      logabs[w] = lo;
    } else {
      sign[w] = v > 0.0 ? 1.0 : -1.0;
      logabs[w] = fmax(log(fabs(v)) + lsc, lo);
    }
  }
}

// ------------------------------------------------------------ environment reuse
// Per-walker cache of the column environments of the exact contraction:
// L_y = columns 0 .. y-1 absorbed from the left, R_y = columns y .. Ly-1
// absorbed from the right (y = 0 .. Ly; both indexed [a][b] at the cut left
// of column y, stride PB), each renormalised with its log scale.  A
// configuration that differs from its walker only in columns c0 .. c1 has
//   psi' = gsign' * twist(L_c0) . C'_c0 ... C'_c1 . R_{c1+1},
// where the twist accounts for the local signs of the sites left of c0:
// S_s carries (-1)^{q(d) P_right(s)}, and P_right(s) of a site in row x,
// column < c0 changes parity exactly when the number of singly occupied
// sites of row x in columns >= c0 does (flip_x).  Flipping P_right on all of
// row x left of the cut multiplies those slices by (-1)^{q(d)}; each slice
// is Z2-even, so (-1)^{q(d)} = (-1)^{q(p) + q(l) + q(u) + q(r)} on it, and
// along the row the horizontal bonds cancel in pairs: the twist of row x
// equals sigma_x (-1)^{q(h_x)} times the twist of row x-1 (its d legs are
// row x's u legs), with sigma_x = (-1)^{singly occupied sites of row x left
// of the cut} and h_x the horizontal bond of row x crossing the cut.  Hence
//   twist = prod_{x: F_x odd} sigma_x (-1)^{q(h_x)},  F_x = sum_{x' >= x} flip_x',
// a sign and a parity on the bond indices of L_c0.  R_{c1+1} needs none
// (P_right looks right only).  L_y depends on the configuration right of
// its cut through those parities, so each L_y keeps the row parities it was
// built with, and flip_x is taken against them: an L_y stays usable after
// accepted moves right of its cut (the sweep relies on this), and only a
// change of the columns left of the cut makes it stale.
__host__ __device__ inline int64_t env_stride(int Ly) {
  const int64_t n = 2 * (int64_t)(Ly + 1) * PB * PB + 3 * (Ly + 1);
  return (n + 31) & ~(int64_t)31;
}
__device__ inline double *env_L(double *e, int y) { return e + (int64_t)y * PB * PB; }
__device__ inline double *env_R(double *e, int Ly, int y) {
  return e + (int64_t)(Ly + 1 + y) * PB * PB;
}
// log scales: L at [y], R at [Ly + 1 + y]; then, at [2 (Ly + 1) + y], the
// row mask of L_y: bit x = parity of the singly occupied sites of row x in
// columns >= y when L_y was built (the P_right parities its slices carry)
__device__ inline double *env_sc(double *e, int Ly) { return e + 2 * (int64_t)(Ly + 1) * PB * PB; }

// bond dimension of a row's horizontal bond at the cut left of column y
__device__ __forceinline__ int cut_dim(const PepsGeo &g, int y) {
  return (y == 0 || y == g.Ly) ? 1 : g.D;
}

// singly occupied sites of row x in columns [ya, yb)
__device__ __forceinline__ int singles(const PepsGeo &g, const int64_t *c, int x, int ya,
                                      int yb) {
  int n = 0;
  for (int y = ya; y < yb; ++y) {
    const int64_t v = c[x * g.Ly + y];
    n += (v == 1 || v == 2);
  }
  return n;
}

// buf <- L_c0 of the environment block e, twisted to the configuration c
// (whose columns left of c0 are those L_c0 was built from): the rows whose
// P_right parities differ from the ones L_c0 carries (F_x odd) contribute
// sigma_x (-1)^{q(h_x)}; the sign is folded into the entries.  Block-wide;
// tw: a shared int.
__device__ void load_left(const PepsGeo &g, const double *e, int c0, const int64_t *c,
                          double *buf, int *tw) {
  const double *sc = e + 2 * (int64_t)(g.Ly + 1) * PB * PB;
  if (threadIdx.x == 0) {
    const int built = (int)sc[2 * (g.Ly + 1) + c0];
    int F = 0, rowsm = 0, sg = 0;
    if (c0 > 0) {
      for (int x = g.Lx - 1; x >= 0; --x) {
        F ^= (singles(g, c, x, c0, g.Ly) ^ (built >> x)) & 1;
        if (F) {
          rowsm |= 1 << x;
          sg ^= singles(g, c, x, 0, c0) & 1;
        }
      }
    }
    *tw = rowsm | (sg << 8);
  }
  __syncthreads();
  const int twr = *tw & 0xFF;
  const bool neg = (*tw >> 8) & 1;
  const int dc = cut_dim(g, c0);
  const int br = g.Lx - g.tr;
  for (int idx = threadIdx.x; idx < PB * PB; idx += PT) {
    const int a = idx / PB, b = idx - a * PB;
    double v = e[(int64_t)c0 * PB * PB + idx];
    int q = neg ? 1 : 0;
    if (twr) {
      // the row components of the cut's bond indices: a = h_0 (dc) + h_1
      // (two top rows) or h_0; b likewise for the bottom rows
      const int ht0 = g.tr == 2 ? a / dc : a, ht1 = g.tr == 2 ? a % dc : 0;
      const int hb0 = br == 2 ? b / dc : b, hb1 = br == 2 ? b % dc : 0;
      if (twr & 1) q += qpar(ht0, dc);
      if (g.tr == 2 && (twr & 2)) q += qpar(ht1, dc);
      if (twr & (1 << g.tr)) q += qpar(hb0, dc);
      if (br == 2 && (twr & (1 << (g.tr + 1)))) q += qpar(hb1, dc);
    }
    buf[idx] = (q & 1) ? -v : v;
  }
  __syncthreads();
}

// L_0 = R_Ly = [[1]], scales 0
__global__ void peps_env_init_kernel(int64_t B, int Ly, double *env) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= B) return;
  double *e = env + w * env_stride(Ly);
  env_L(e, 0)[0] = 1.0;
  env_R(e, Ly, Ly)[0] = 1.0;
  env_sc(e, Ly)[0] = 0.0;
  env_sc(e, Ly)[Ly + 1 + Ly] = 0.0;
  env_sc(e, Ly)[2 * (Ly + 1)] = 0.0;   // L_0 holds no site: no twist applies
}

// dir 0: L_{y+1} from L_y for y = y0 .. y1-1; dir 1: R_y from R_{y+1} for
// y = y1-1 down to y0 (one workgroup per walker, the walker's columns)
__global__ void __launch_bounds__(PT)
    peps_env_kernel(PepsArgs A, int64_t B, const int64_t *__restrict__ cfg, int y0, int y1,
                    int dir, double *env) {
  const int64_t w = blockIdx.x;
  if (w >= B) return;
  __shared__ CfgSh cs;
  __shared__ ColSh col;
  __shared__ double buf[2][PB * PB];
  __shared__ double red[PT / 64];
  const PepsGeo &g = A.g;
  load_cfg(g, cfg + w * g.N, cs);
  double *e = env + w * env_stride(g.Ly);
  double *sc = env_sc(e, g.Ly);
  __shared__ int tw;
  int cur = 0;
  if (dir == 0) {
    // L_y0 may predate accepted moves right of its cut: bring its signs to
    // the walker's present configuration before absorbing further columns
    load_left(g, e, y0, cfg + w * g.N, buf[0], &tw);
    double lsc = sc[y0];
    for (int y = y0; y < y1; ++y) {
      build_column(g, cs, y, A.dmap, A.theta, col);
      absorb_left(col, buf[cur], buf[cur ^ 1]);
      cur ^= 1;
      lsc += renorm(buf[cur], col.da2, col.db2, red);
      for (int idx = threadIdx.x; idx < PB * PB; idx += PT) env_L(e, y + 1)[idx] = buf[cur][idx];
      if (threadIdx.x == 0) {
        sc[y + 1] = lsc;
        int mask = 0;
        for (int x = 0; x < g.Lx; ++x) mask |= (singles(g, cfg + w * g.N, x, y + 1, g.Ly) & 1) << x;
        sc[2 * (g.Ly + 1) + y + 1] = (double)mask;
      }
    }
  } else {
    for (int idx = threadIdx.x; idx < PB * PB; idx += PT) buf[0][idx] = env_R(e, g.Ly, y1)[idx];
    double lsc = sc[g.Ly + 1 + y1];
    __syncthreads();
    for (int y = y1 - 1; y >= y0; --y) {
      build_column(g, cs, y, A.dmap, A.theta, col);
      absorb_right(col, buf[cur], buf[cur ^ 1]);
      cur ^= 1;
      lsc += renorm(buf[cur], col.da, col.db, red);
      for (int idx = threadIdx.x; idx < PB * PB; idx += PT)
        env_R(e, g.Ly, y)[idx] = buf[cur][idx];
      if (threadIdx.x == 0) sc[g.Ly + 1 + y] = lsc;
    }
  }
}

// Amplitudes of `rows` configurations, row k equal to walker parent[k]
// (parent NULL: walker k) outside columns c0 .. c1: from the walker's cached
// L_c0 (twisted to row k's configuration) and R_{c1+1}, absorbing only
// columns c0 .. c1.  cmove
// (may be NULL): per row the hop from_orbital | to_orbital << 8 (orbital =
// site, or N + site for down), c0 / c1 its columns; otherwise c0 / c1 as
// given for every row.  flag (may be NULL): rows with !(flag & 1) skipped.
__global__ void __launch_bounds__(PT)
    peps_amp_env_kernel(PepsArgs A, int64_t rows, const int64_t *__restrict__ cfg_rows,
                        const int32_t *__restrict__ parent, const uint32_t *__restrict__ cmove,
                        int c0u, int c1u,
                        const int32_t *__restrict__ flag, const double *__restrict__ env,
                        double *sign, double *logabs) {
  const int64_t k = blockIdx.x;
  if (k >= rows) return;
  if (flag && !(flag[k] & 1)) return;
  __shared__ CfgSh cs;
  __shared__ ColSh col;
  __shared__ double buf[2][PB * PB];
  __shared__ double red[PT / 64];
  __shared__ int tw;
  const PepsGeo &g = A.g;
  const int64_t w = parent ? (int64_t)parent[k] : k;
  const int64_t *cn = cfg_rows + k * g.N;
  load_cfg(g, cn, cs);
  int c0 = c0u, c1 = c1u;
  if (cmove) {
    const uint32_t mv = cmove[k];
    int sa = mv & 0xFF, sb = (mv >> 8) & 0xFF;
    if (sa >= g.N) sa -= g.N;
    if (sb >= g.N) sb -= g.N;
    const int ya = sa % g.Ly, yb = sb % g.Ly;
    c0 = ya < yb ? ya : yb;
    c1 = ya < yb ? yb : ya;
  }
  const double *e = env + w * env_stride(g.Ly);
  const double *sc = e + 2 * (int64_t)(g.Ly + 1) * PB * PB;
  load_left(g, e, c0, cn, buf[0], &tw);
  double lsc = sc[c0];
  int cur = 0;
  for (int y = c0; y <= c1; ++y) {
    build_column(g, cs, y, A.dmap, A.theta, col);
    absorb_left(col, buf[cur], buf[cur ^ 1]);
    cur ^= 1;
    lsc += renorm(buf[cur], col.da2, col.db2, red);
  }
  // psi' / scales = sum_{a,b} L'[a][b] R_{c1+1}[a][b]
  const double *Rr = e + (int64_t)(g.Ly + 1 + c1 + 1) * PB * PB;
  const int dr_ = cut_dim(g, c1 + 1);
  const int br = g.Lx - g.tr;
  const int ea = g.tr == 2 ? dr_ * dr_ : dr_, eb = br == 2 ? dr_ * dr_ : dr_;
  double part = 0.0;
  for (int idx = threadIdx.x; idx < PB * PB; idx += PT) {
    const int a = idx / PB, b = idx - a * PB;
    if (a < ea && b < eb) part += buf[cur][idx] * Rr[idx];
  }
  for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    double dot = 0.0;
    for (int q = 0; q < PT / 64; ++q) dot += red[q];
    const double v = dot * cs.gsign;
    const double lo = log(1e-45);
    if (v == 0.0 || !std::isfinite(v)) {
      sign[k] = 0.0;
      logabs[k] = lo;
    } else {
      sign[k] = v > 0.0 ? 1.0 : -1.0;
      logabs[k] = fmax(log(fabs(v)) + lsc + sc[g.Ly + 1 + c1 + 1], lo);
    }
  }
}

// ------------------------------------------------------------ log-derivatives
// G[row][k] = d ln|psi| / d theta_k for the tensor entries (zero for entries
// of local states other than x_s); sign / logabs of each configuration.
__global__ void __launch_bounds__(PT)
    peps_grad_kernel(PepsArgs A, int64_t rows, const int64_t *__restrict__ cfg,
                     double *G, int64_t ldg, double *sign, double *logabs) {
  const int64_t w = blockIdx.x;
  if (w >= rows) return;
  __shared__ CfgSh cs;
  __shared__ ColSh col;
  __shared__ double Lenv[PMAX_LY + 1][PB * PB];
  __shared__ double Lsc[PMAX_LY + 1];
  __shared__ double R[2][PB * PB];
  __shared__ double dTop[PB * PMAX_D * PB], dBot[PB * PMAX_D * PB];
  __shared__ double dsl[PMAX_LX][SLICE];
  __shared__ double red[PT / 64];
  const PepsGeo &g = A.g;
  load_cfg(g, cfg + w * g.N, cs);
  const double *th = A.theta + w * A.theta_stride;
  double *grow = G + w * ldg;
  for (int64_t k = threadIdx.x; k < A.n_params; k += PT) grow[k] = 0.0;
  if (threadIdx.x == 0) {
    Lenv[0][0] = 1.0;
    Lsc[0] = 0.0;
  }
  __syncthreads();
  for (int y = 0; y < g.Ly; ++y) {
    build_column(g, cs, y, A.dmap, th, col);
    absorb_left(col, Lenv[y], Lenv[y + 1]);
    const double l = renorm(Lenv[y + 1], col.da2, col.db2, red);
    if (threadIdx.x == 0) Lsc[y + 1] = Lsc[y] + l;
    __syncthreads();
  }
  const double fin = Lenv[g.Ly][0];
  const double lpsi = Lsc[g.Ly];
  if (threadIdx.x == 0) {
    const double v = fin * cs.gsign;
    const double lo = log(1e-45);
    if (v == 0.0 || !std::isfinite(v)) {
      sign[w] = 0.0;
      logabs[w] = lo;
    } else {
      sign[w] = v > 0.0 ? 1.0 : -1.0;
      logabs[w] = fmax(log(fabs(v)) + lpsi, lo);
    }
  }
  if (fin == 0.0 || !std::isfinite(fin)) return;   // zero amplitude: O row stays 0
  // right sweep
  if (threadIdx.x == 0) R[0][0] = 1.0;
  __syncthreads();
  int cur = 0;
  double rsc = 0.0;
  const int D = g.D;
  const int br = g.Lx - g.tr;
  for (int y = g.Ly - 1; y >= 0; --y) {
    build_column(g, cs, y, A.dmap, th, col);
    const double *Ly_ = Lenv[y];
    const double *Rr = R[cur];
    const int da = col.da, da2 = col.da2, db = col.db, db2 = col.db2, dv = col.dv;
    // d psi / d Top[a][v][a'] = sum_{b,b'} L[a][b] Bot[b][v][b'] R[a'][b']
    // X[a][v][b'] = sum_b L[a][b] Bot[b][v][b']
    for (int idx = threadIdx.x; idx < da * dv * db2; idx += PT) {
      const int b2 = idx % db2, v = (idx / db2) % dv, a = idx / (db2 * dv);
      double acc = 0.0;
      for (int b = 0; b < db; ++b) acc += Ly_[a * PB + b] * col.bot[(b * dv + v) * db2 + b2];
      col.X[idx] = acc;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < da * dv * da2; idx += PT) {
      const int a2 = idx % da2, v = (idx / da2) % dv, a = idx / (da2 * dv);
      double acc = 0.0;
      for (int b2 = 0; b2 < db2; ++b2) acc += col.X[(a * dv + v) * db2 + b2] * Rr[a2 * PB + b2];
      dTop[idx] = acc;
    }
    __syncthreads();
    // d psi / d Bot[b][v][b'] = sum_{a,a'} L[a][b] Top[a][v][a'] R[a'][b']
    for (int idx = threadIdx.x; idx < db * dv * da2; idx += PT) {
      const int a2 = idx % da2, v = (idx / da2) % dv, b = idx / (da2 * dv);
      double acc = 0.0;
      for (int a = 0; a < da; ++a) acc += Ly_[a * PB + b] * col.top[(a * dv + v) * da2 + a2];
      col.X[idx] = acc;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < db * dv * db2; idx += PT) {
      const int b2 = idx % db2, v = (idx / db2) % dv, b = idx / (db2 * dv);
      double acc = 0.0;
      for (int a2 = 0; a2 < da2; ++a2) acc += col.X[(b * dv + v) * da2 + a2] * Rr[a2 * PB + b2];
      dBot[idx] = acc;
    }
    __syncthreads();
    // rows of each half
    const int dlc = (y == 0) ? 1 : D, drc = (y == g.Ly - 1) ? 1 : D;
    if (g.tr == 1) {
      for (int idx = threadIdx.x; idx < dlc * dv * drc; idx += PT) dsl[0][idx] = dTop[idx];
    } else {
      // dA[l0][m][r0] (row 0: [l0][0][m][r0])
      for (int idx = threadIdx.x; idx < dlc * D * drc; idx += PT) {
        const int r0 = idx % drc, m = (idx / drc) % D, l0 = idx / (drc * D);
        double acc = 0.0;
        for (int l1 = 0; l1 < dlc; ++l1)
          for (int v = 0; v < dv; ++v)
            for (int r1 = 0; r1 < drc; ++r1)
              acc += dTop[((l0 * dlc + l1) * dv + v) * da2 + r0 * drc + r1] *
                     col.sl[1][((l1 * D + m) * dv + v) * drc + r1];
        dsl[0][idx] = acc;
      }
      // dB[l1][m][v][r1]
      for (int idx = threadIdx.x; idx < dlc * D * dv * drc; idx += PT) {
        const int r1 = idx % drc, v = (idx / drc) % dv, m = (idx / (drc * dv)) % D,
                  l1 = idx / (drc * dv * D);
        double acc = 0.0;
        for (int l0 = 0; l0 < dlc; ++l0)
          for (int r0 = 0; r0 < drc; ++r0)
            acc += dTop[((l0 * dlc + l1) * dv + v) * da2 + r0 * drc + r1] *
                   col.sl[0][(l0 * D + m) * drc + r0];
        dsl[1][idx] = acc;
      }
    }
    const int xb = g.tr;
    if (br == 1) {
      for (int idx = threadIdx.x; idx < dlc * dv * drc; idx += PT) dsl[xb][idx] = dBot[idx];
    } else {
      // dC[l2][v][m][r2] (row xb), dE[l3][m][0][r3] (row xb+1)
      for (int idx = threadIdx.x; idx < dlc * dv * D * drc; idx += PT) {
        const int r2 = idx % drc, m = (idx / drc) % D, v = (idx / (drc * D)) % dv,
                  l2 = idx / (drc * D * dv);
        double acc = 0.0;
        for (int l3 = 0; l3 < dlc; ++l3)
          for (int r3 = 0; r3 < drc; ++r3)
            acc += dBot[((l2 * dlc + l3) * dv + v) * db2 + r2 * drc + r3] *
                   col.sl[xb + 1][(l3 * D + m) * drc + r3];
        dsl[xb][idx] = acc;
      }
      for (int idx = threadIdx.x; idx < dlc * D * drc; idx += PT) {
        const int r3 = idx % drc, m = (idx / drc) % D, l3 = idx / (drc * D);
        double acc = 0.0;
        for (int l2 = 0; l2 < dlc; ++l2)
          for (int v = 0; v < dv; ++v)
            for (int r2 = 0; r2 < drc; ++r2)
              acc += dBot[((l2 * dlc + l3) * dv + v) * db2 + r2 * drc + r3] *
                     col.sl[xb][((l2 * dv + v) * D + m) * drc + r2];
        dsl[xb + 1][idx] = acc;
      }
    }
    __syncthreads();
    // scatter: d ln|psi| / d T = d psi / d S * sign / psi
    const double f = exp(Lsc[y] + rsc - lpsi) / fin;
    for (int x = 0; x < g.Lx; ++x) {
      int dl, du, dd, dr;
      site_dims(g, x, y, dl, du, dd, dr);
      const int s = x * g.Ly + y;
      const int n = dl * du * dd * dr;
      const int base = cs.doff[s] + cs.p[s] * n;
      const int pr = cs.pr[s];
      for (int idx = threadIdx.x; idx < n; idx += PT) {
        const int r = idx % dr, d = (idx / dr) % dd;
        const int k = A.dmap[base + idx];
        if (k < 0) continue;
        double v = dsl[x][idx] * f;
        if ((qpar(d, dd) * (qpar(r, dr) + pr)) & 1) v = -v;
        grow[k] = v;
      }
    }
    // right environment for the next column to the left
    absorb_right(col, Rr, R[cur ^ 1]);
    cur ^= 1;
    rsc += renorm(R[cur], col.da, col.db, red);
  }
}

// ------------------------------------------------------------ NN-fTNS MLP
namespace {

__device__ __forceinline__ double act_d(double z, int act) {
  return act == 0 ? (z > 0.0 ? z : 0.0) : tanh(z);
}

__device__ __forceinline__ double dact_d(double z, int act) {
  if (act == 0) return z > 0.0 ? 1.0 : 0.0;
  const double t = tanh(z);
  return 1.0 - t * t;
}

}  // namespace

// z1[row][u] = b1[u] + sum_{occupied orbitals o} W1[u][o]; h1 = act(z1)
__global__ void __launch_bounds__(256)
    pnn_layer1_kernel(int64_t rows, int N, int H, int act, const int64_t *__restrict__ cfg,
                      const double *__restrict__ W1, const double *__restrict__ b1,
                      double *z1, double *h1) {
  const int64_t w = blockIdx.x;
  if (w >= rows) return;
  __shared__ int occ[2 * PMAX_NS];
  __shared__ int nocc;
  if (threadIdx.x == 0) {
    int k = 0;
    const int64_t *c = cfg + w * N;
    for (int i = 0; i < N; ++i)
      if (c[i] & 2) occ[k++] = i;
    for (int i = 0; i < N; ++i)
      if (c[i] & 1) occ[k++] = N + i;
    nocc = k;
  }
  __syncthreads();
  const int n_orb = 2 * N;
  for (int u = threadIdx.x; u < H; u += blockDim.x) {
    double z = b1[u];
    for (int k = 0; k < nocc; ++k) z += W1[(int64_t)u * n_orb + occ[k]];
    z1[w * H + u] = z;
    h1[w * H + u] = act_d(z, act);
  }
}

// z += b ; h = act(z)
__global__ void pnn_bias_act_kernel(int64_t n, int H, int act, double *z,
                                    const double *__restrict__ b, double *h) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = z[i] + b[i % H];
    z[i] = v;
    h[i] = act_d(v, act);
  }
}

// out[row][k] = theta[k] + eta * (out[row][k] + bo[k])
__global__ void pnn_theta_kernel(int64_t n, int64_t P, double eta, const double *theta,
                                 const double *__restrict__ bo, double *out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = i % P;
    out[i] = theta[k] + eta * (out[i] + bo[k]);
  }
}

// multiplicative correction: y[row][k] += bo[k] (the network output, kept for
// the backward pass) and out[row][k] = theta[k] (1 + eta y[row][k])
__global__ void pnn_theta_mul_kernel(int64_t n, int64_t P, double eta, const double *theta,
                                     const double *__restrict__ bo, double *y, double *out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = i % P;
    const double v = y[i] + bo[k];
    y[i] = v;
    out[i] = theta[k] * (1.0 + eta * v);
  }
}

// multiplicative correction, backward: from G = d ln|psi| / d theta_eff,
// y <- G (1 + eta y) (the theta block of O) and G <- G theta (the upstream of
// the network, whose output enters as eta theta y)
__global__ void pnn_mul_grad_kernel(int64_t n, int64_t P, double eta, const double *theta,
                                    double *y, double *G) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = i % P;
    const double g = G[i];
    y[i] = g * (1.0 + eta * y[i]);
    G[i] = g * theta[k];
  }
}

// dz[i] = dh[i] * act'(z[i])
__global__ void pnn_dact_kernel(int64_t n, int act, const double *z, double *dh) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dh[i] *= dact_d(z[i], act);
}

// O row: [gth (P) | W1 (H x 2N) | b1 | W2 (H x H) | b2 | Wo (P x H) | bo (P)]
// with dOut = eta g; one block per sample.  FACT: the factors of that row
// instead, [gth | n (2N) | dz1 | h1 | dz2 | h2] (W1 = dz1 n^T, W2 = dz2 h1^T,
// Wo = eta g h2^T, the biases dz1, dz2, eta g), followed by g itself when
// gth != g (the multiplicative correction).  Additive correction: gth = g = G.
template <typename TO, bool FACT = false>
__global__ void __launch_bounds__(256)
    pnn_write_kernel(int64_t rows, int N, int H, int64_t P, int act, double eta,
                     const int64_t *__restrict__ cfg, const double *__restrict__ G,
                     const double *__restrict__ Gth,
                     const double *__restrict__ z1, const double *__restrict__ z2,
                     const double *__restrict__ dz1, const double *__restrict__ dz2,
                     TO *O, int64_t ld) {
  const int64_t w = blockIdx.x;
  if (w >= rows) return;
  extern __shared__ double dsm[];
  double *h1 = dsm, *h2 = h1 + H, *d1 = h2 + H, *d2 = d1 + H;
  __shared__ float nvec[2 * PMAX_NS];
  const int n_orb = 2 * N;
  const int64_t *c = cfg + w * N;
  for (int o = threadIdx.x; o < n_orb; o += blockDim.x)
    nvec[o] = (o < N) ? ((c[o] & 2) ? 1.0f : 0.0f) : ((c[o - N] & 1) ? 1.0f : 0.0f);
  for (int u = threadIdx.x; u < H; u += blockDim.x) {
    h1[u] = act_d(z1[w * H + u], act);
    h2[u] = act_d(z2[w * H + u], act);
    d1[u] = dz1[w * H + u];
    d2[u] = dz2[w * H + u];
  }
  __syncthreads();
  const double *g = G + w * P;
  const double *gth = Gth + w * P;
  TO *row = O + w * ld;
  for (int64_t k = threadIdx.x; k < P; k += blockDim.x) row[k] = (TO)gth[k];
  if (FACT) {
    TO *f = row + P;
    for (int o = threadIdx.x; o < n_orb; o += blockDim.x) f[o] = (TO)nvec[o];
    f += n_orb;
    for (int u = threadIdx.x; u < H; u += blockDim.x) {
      f[u] = (TO)d1[u];
      f[H + u] = (TO)h1[u];
      f[2 * H + u] = (TO)d2[u];
      f[3 * H + u] = (TO)h2[u];
    }
    if (Gth != G) {
      f += 4 * H;
      for (int64_t k = threadIdx.x; k < P; k += blockDim.x) f[k] = (TO)g[k];
    }
    return;
  }
  int64_t off = P;
  for (int64_t idx = threadIdx.x; idx < (int64_t)H * n_orb; idx += blockDim.x) {
    const int u = (int)(idx / n_orb), o = (int)(idx - (int64_t)u * n_orb);
    row[off + idx] = (TO)(d1[u] * nvec[o]);
  }
  off += (int64_t)H * n_orb;
  for (int u = threadIdx.x; u < H; u += blockDim.x) row[off + u] = (TO)d1[u];
  off += H;
  for (int64_t idx = threadIdx.x; idx < (int64_t)H * H; idx += blockDim.x) {
    const int u = (int)(idx / H), i = (int)(idx - (int64_t)u * H);
    row[off + idx] = (TO)(d2[u] * h1[i]);
  }
  off += (int64_t)H * H;
  for (int u = threadIdx.x; u < H; u += blockDim.x) row[off + u] = (TO)d2[u];
  off += H;
  for (int64_t idx = threadIdx.x; idx < P * H; idx += blockDim.x) {
    const int64_t k = idx / H;
    const int u = (int)(idx - k * H);
    row[off + idx] = (TO)(eta * g[k] * h2[u]);
  }
  off += P * H;
  for (int64_t k = threadIdx.x; k < P; k += blockDim.x) row[off + k] = (TO)(eta * g[k]);
}

// FP32 O rows from the FP64 G rows (plain fPEPS)
__global__ void peps_copy_g_kernel(int64_t rows, int64_t P, const double *__restrict__ G,
                                   float *O, int64_t ld) {
  const int64_t w = blockIdx.x;
  if (w >= rows) return;
  for (int64_t k = threadIdx.x; k < P; k += blockDim.x) O[w * ld + k] = (float)G[w * P + k];
}

// proposed configurations: copy of cfg with (i, j) <- (ni, nj) where flagged
__global__ void peps_prop_cfg_kernel(int64_t B, int N, int i, int j,
                                     const int64_t *__restrict__ cfg,
                                     const int32_t *__restrict__ prop, int64_t *pcfg) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * N) return;
  const int64_t w = idx / N;
  const int s = (int)(idx - w * N);
  int64_t v = cfg[idx];
  const int p = prop[w];
  if (p & 1) {
    if (s == i) v = (p >> 8) & 0xF;
    if (s == j) v = (p >> 12) & 0xF;
  }
  pcfg[idx] = v;
}

// Metropolis test with |psi'|^2 / |psi|^2 and commit
// (acc, when given: 1 for the walkers that accepted, else 0)
__global__ void peps_accept_kernel(int64_t B, int N, int i, int j, int64_t *cfg,
                                   const int32_t *__restrict__ prop,
                                   const double *__restrict__ u_acc,
                                   const double *__restrict__ psign,
                                   const double *__restrict__ plog, double *sign,
                                   double *logabs, int64_t *n_accept, int32_t *acc) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= B) return;
  if (acc) acc[w] = 0;
  const int p = prop[w];
  if (!(p & 1)) return;
  const double ratio = exp(2.0 * (plog[w] - logabs[w]));
  if (!(u_acc[w] < ratio)) return;
  if (acc) acc[w] = 1;
  cfg[w * N + i] = (p >> 8) & 0xF;
  cfg[w * N + j] = (p >> 12) & 0xF;
  sign[w] = psign[w];
  logabs[w] = plog[w];
  if (n_accept) n_accept[w] += 1;
}

// ============================================================ host side
namespace {

struct PepsWs {
  double *theta_eff, *y, *z1, *z2, *h1, *h2, *G, *d1, *d2;
  int32_t *prop;
  uint32_t *moves;
  double *uacc, *psign, *plog;
  int64_t *pcfg;
  int64_t *counts, *offsets, *total, *etas;
  double *coeffs, *csign, *clog;
  // FP32 amplitudes of a shared parameter vector: the signed slice table
  float *tbl;
  // chi-truncated boundary MPS: slots of bm_pass configurations (bm_gpass
  // for the log-derivative's per-row environments); the sweep's boundary
  // cache of the B walkers (plain fPEPS)
  void *bm, *bmc;
  int64_t bm_pass, bm_gpass;
  // environment reuse: the walkers' column environments, and the walker and
  // hop of each connected configuration
  double *env;
  int32_t *cparent;
  uint32_t *cmove;
  size_t bytes;
};

size_t alb(size_t x) { return (x + 255) & ~(size_t)255; }

// configurations per pass of the boundary-MPS kernels (their workspace)
constexpr int64_t BM_PASS = 2048, BM_GPASS = 1024;

// the chi-truncated boundary MPS instead of the exact column contraction:
// chi > 0 and either the lattice is beyond the exact kernels or chi is below
// the largest bond of a side (D^rows)
bool use_bmps(const vmcx_peps_model &m) {
  if (m.chi <= 0) return false;
  if (m.Lx > PMAX_LX || m.Ly > PMAX_LY || m.Lx * m.Ly > PMAX_N) return true;
  int64_t big = 1;
  for (int i = 0; i < m.Lx - m.Lx / 2; ++i) big *= m.D;
  return m.chi < big;
}

PepsWs peps_layout(const vmcx_peps_model &m, int64_t B, int64_t chunk, int n_edges,
                   char *base, bool reuse = false) {
  PepsWs L{};
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t o = off;
    off += alb(b);
    return base ? (void *)(base + o) : nullptr;
  };
  const int64_t P = m.n_params, H = m.hidden > 0 ? m.hidden : 0;
  const int N = m.Lx * m.Ly;
  const int64_t rows = chunk > B ? chunk : B;   // the sweep evaluates B at once
  L.theta_eff = (double *)take(H ? (size_t)rows * P * 8 : 0);
  // the network output of the multiplicative correction (kept for its backward)
  L.y = (double *)take(H && m.mode == 1 ? (size_t)rows * P * 8 : 0);
  L.z1 = (double *)take((size_t)rows * H * 8);
  L.z2 = (double *)take((size_t)rows * H * 8);
  L.h1 = (double *)take((size_t)rows * H * 8);
  L.h2 = (double *)take((size_t)rows * H * 8);
  L.d1 = (double *)take((size_t)rows * H * 8);
  L.d2 = (double *)take((size_t)rows * H * 8);
  L.G = (double *)take((size_t)chunk * P * 8);
  L.prop = (int32_t *)take((size_t)B * 4);
  L.moves = (uint32_t *)take((size_t)B * 4);
  L.uacc = (double *)take((size_t)B * 8);
  L.psign = (double *)take((size_t)B * 8);
  L.plog = (double *)take((size_t)B * 8);
  L.pcfg = (int64_t *)take((size_t)B * N * 8);
  const int64_t T = B * 2 * (int64_t)n_edges;
  L.counts = (int64_t *)take((size_t)B * 8);
  L.offsets = (int64_t *)take((size_t)B * 8);
  L.total = (int64_t *)take(8);
  L.etas = (int64_t *)take((size_t)T * N * 8);
  L.coeffs = (double *)take((size_t)T * 8);
  L.csign = (double *)take((size_t)T * 8);
  L.clog = (double *)take((size_t)T * 8);
  L.tbl = (float *)take(m.amp_fp32 && H == 0 ? (size_t)N * 8 * 256 * 4 : 0);
  if (use_bmps(m)) {
    L.bm_pass = rows < BM_PASS ? (rows > 0 ? rows : 1) : BM_PASS;
    L.bm_gpass = chunk < BM_GPASS ? (chunk > 0 ? chunk : 1) : BM_GPASS;
    const int64_t ba = bmps_workspace_bytes(m.Lx, m.Ly, m.D, m.chi, L.bm_pass);
    const int64_t bg = bmps_logderiv_workspace_bytes(m.Lx, m.Ly, m.D, m.chi, L.bm_gpass);
    L.bm = take((size_t)(ba > bg ? ba : bg));
    if (H == 0 && B > 0) L.bmc = take((size_t)bmps_sweep_cache_bytes(m.Lx, m.Ly, m.D, m.chi, B));
  }
  if (reuse) L.env = (double *)take((size_t)B * env_stride(m.Ly) * 8);
  if (reuse || (use_bmps(m) && H == 0 && B > 0)) {
    // the walker and hop of each connected configuration (environment reuse,
    // and the boundary cache of a chi-truncated plain fPEPS)
    L.cparent = (int32_t *)take((size_t)T * 4);
    L.cmove = (uint32_t *)take((size_t)T * 4);
  }
  L.bytes = off;
  return L;
}

int check_peps(const vmcx_peps_model *m) {
  VMCX_CHECK_ARG(m != nullptr, "null model");
  VMCX_CHECK_ARG(m->D >= 1 && m->D <= PMAX_D, "D must be 1..4");
  if (use_bmps(*m)) {
    const char *why = bmps_check(m->Lx, m->Ly, m->D, m->chi);
    VMCX_CHECK_ARG(why == nullptr, why ? why : "");
    VMCX_CHECK_ARG(!m->amp_fp32,
                   "FP32 amplitudes need the exact contraction (Lx <= 4, chi <= 0 or chi >= "
                   "D^rows)");
  } else {
    VMCX_CHECK_ARG(m->Lx >= 2 && m->Lx <= PMAX_LX,
                   "Lx must be 2..4 for the exact contraction (chi > 0 selects the boundary MPS)");
    VMCX_CHECK_ARG(m->Ly >= 1 && m->Ly <= PMAX_LY, "Ly must be 1..8 for the exact contraction");
  }
  VMCX_CHECK_ARG(m->n_params > 0 && m->dense_map && m->theta, "missing parameters");
  VMCX_CHECK_ARG(m->hidden >= 0 && m->hidden <= 1024, "hidden must be 0..1024");
  VMCX_CHECK_ARG(m->act == 0 || m->act == 1, "act must be 0 (ReLU) or 1 (tanh)");
  VMCX_CHECK_ARG(m->mode == 0 || m->mode == 1, "mode must be 0 (additive) or 1 (multiplicative)");
  VMCX_CHECK_ARG(m->mode == 0 || m->hidden > 0, "the multiplicative correction needs hidden > 0");
  if (m->hidden > 0)
    VMCX_CHECK_ARG(m->W1 && m->b1 && m->W2 && m->b2 && m->Wo && m->bo, "missing MLP weights");
  return 0;
}

PepsArgs args_of(const vmcx_peps_model &m, const double *theta, int64_t stride) {
  PepsArgs A;
  A.g.Lx = m.Lx;
  A.g.Ly = m.Ly;
  A.g.D = m.D;
  A.g.N = m.Lx * m.Ly;
  A.g.tr = m.Lx / 2;
  A.dmap = m.dense_map;
  A.theta = theta;
  A.theta_stride = stride;
  A.n_params = m.n_params;
  return A;
}

#define PEPS_BLAS(call)                                                       \
  do {                                                                       \
    rocblas_status s_ = (call);                                              \
    if (s_ != rocblas_status_success) {                                      \
      ::vmcx::set_error(std::string(__func__) + ": " #call " -> " +          \
                        rocblas_status_to_string(s_));                       \
      return 3;                                                              \
    }                                                                        \
  } while (0)

// theta(x) rows for `rows` configurations (NN-fTNS); keeps z1 / z2
int mlp_theta(const vmcx_peps_model &m, const PepsWs &L, int64_t rows, const int64_t *cfg,
              hipStream_t st) {
  const int H = m.hidden, N = m.Lx * m.Ly;
  const int64_t P = m.n_params;
  rocblas_handle h = blas_handle_for(st);
  VMCX_CHECK_ARG(h != nullptr, "rocblas handle");
  hipLaunchKernelGGL(pnn_layer1_kernel, dim3((unsigned)rows), dim3(256), 0, st, rows, N, H,
                     m.act, cfg, m.W1, m.b1, L.z1, L.h1);
  VMCX_LAUNCHED();
  const double one = 1.0, zero = 0.0;
  // column-major: z2 (H x rows) = W2^T(col-major view of row-major W2) h1
  PEPS_BLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, H,
                          (rocblas_int)rows, H, &one, m.W2, H, L.h1, H, &zero, L.z2, H));
  const int64_t nh = rows * H;
  const unsigned gh = (unsigned)((nh + 255) / 256 < 8192 ? (nh + 255) / 256 : 8192);
  hipLaunchKernelGGL(pnn_bias_act_kernel, dim3(gh), dim3(256), 0, st, nh, H, m.act, L.z2, m.b2,
                     L.h2);
  VMCX_LAUNCHED();
  double *out = m.mode == 1 ? L.y : L.theta_eff;
  PEPS_BLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none,
                          (rocblas_int)P, (rocblas_int)rows, H, &one, m.Wo, H, L.h2, H, &zero,
                          out, (rocblas_int)P));
  const int64_t nt = rows * P;
  const unsigned gt = (unsigned)((nt + 255) / 256 < 16384 ? (nt + 255) / 256 : 16384);
  if (m.mode == 1)
    hipLaunchKernelGGL(pnn_theta_mul_kernel, dim3(gt), dim3(256), 0, st, nt, P, m.eta, m.theta,
                       m.bo, L.y, L.theta_eff);
  else
    hipLaunchKernelGGL(pnn_theta_kernel, dim3(gt), dim3(256), 0, st, nt, P, m.eta, m.theta,
                       m.bo, L.theta_eff);
  VMCX_LAUNCHED();
  return 0;
}

// amplitudes of `rows` configurations (rows <= chunk for NN-fTNS)
int amp_rows(const vmcx_peps_model &m, const PepsWs &L, int64_t rows, const int64_t *cfg,
             const int32_t *flag, double *sign, double *logabs, hipStream_t st) {
  if (rows == 0) return 0;
  PepsArgs A = args_of(m, m.theta, 0);
  if (m.hidden > 0) {
    const int rc = mlp_theta(m, L, rows, cfg, st);
    if (rc) return rc;
    A = args_of(m, L.theta_eff, m.n_params);
  }
  if (use_bmps(m))
    return bmps_amplitudes(m.Lx, m.Ly, m.D, m.chi, A.dmap, A.theta, A.theta_stride, rows, cfg,
                           flag, sign, logabs, L.bm, L.bm_pass, st);
  if (m.amp_fp32) {
    const float *tbl = nullptr;
    if (m.hidden == 0) {
      const int64_t ne = (int64_t)A.g.N * 8 * 256;
      hipLaunchKernelGGL(peps_slice_table_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256),
                         0, st, A, L.tbl);
      VMCX_LAUNCHED();
      tbl = L.tbl;
    }
    const dim3 grid((unsigned)((rows + P32_W - 1) / P32_W)), blk(64 * P32_W);
    switch (m.D) {
      case 1: hipLaunchKernelGGL(peps_amp32_kernel<1>, grid, blk, 0, st, A, rows, cfg, flag, tbl, sign, logabs); break;
      case 2: hipLaunchKernelGGL(peps_amp32_kernel<2>, grid, blk, 0, st, A, rows, cfg, flag, tbl, sign, logabs); break;
      case 3: hipLaunchKernelGGL(peps_amp32_kernel<3>, grid, blk, 0, st, A, rows, cfg, flag, tbl, sign, logabs); break;
      default: hipLaunchKernelGGL(peps_amp32_kernel<4>, grid, blk, 0, st, A, rows, cfg, flag, tbl, sign, logabs); break;
    }
  } else
    hipLaunchKernelGGL(peps_amp_kernel, dim3((unsigned)rows), dim3(PT), 0, st, A, rows, cfg,
                       flag, sign, logabs);
  VMCX_LAUNCHED();
  return 0;
}

}  // namespace

int64_t peps_dense_size(const vmcx_peps_model *m) {
  if (!m || m->Lx < 1 || m->Ly < 1 || m->D < 1) return -1;
  int64_t n = 0;
  for (int x = 0; x < m->Lx; ++x)
    for (int y = 0; y < m->Ly; ++y) {
      const int dl = (y == 0) ? 1 : m->D, du = (x == 0) ? 1 : m->D;
      const int dd = (x == m->Lx - 1) ? 1 : m->D, dr = (y == m->Ly - 1) ? 1 : m->D;
      n += 4 * dl * du * dd * dr;
    }
  return n;
}

int64_t peps_workspace_bytes(const vmcx_peps_model *m, int64_t B, int64_t chunk, int n_edges) {
  if (check_peps(m) || B < 0 || chunk < 1 || n_edges < 0) return -1;
  return (int64_t)peps_layout(*m, B, chunk, n_edges, nullptr).bytes;
}

int64_t peps_reuse_workspace_bytes(const vmcx_peps_model *m, int64_t B, int64_t chunk,
                                   int n_edges) {
  if (check_peps(m) || B < 0 || chunk < 1 || n_edges < 0) return -1;
  return (int64_t)peps_layout(*m, B, chunk, n_edges, nullptr, true).bytes;
}

namespace {

int check_reuse(const vmcx_peps_model *m) {
  if (check_peps(m)) return 1;
  VMCX_CHECK_ARG(!use_bmps(*m), "environment reuse needs the exact contraction (chi <= 0)");
  VMCX_CHECK_ARG(m->hidden == 0,
                 "environment reuse needs site tensors that do not depend on the "
                 "configuration (plain fPEPS, hidden = 0)");
  return 0;
}

// column environments of all walkers: L_{y0+1} .. L_{y1} (dir 0) or R_{y1-1}
// .. R_{y0} (dir 1)
int env_sweep(const vmcx_peps_model &m, const PepsWs &L, int64_t B, const int64_t *cfg, int y0,
              int y1, int dir, hipStream_t st) {
  if (y1 <= y0 || B == 0) return 0;
  const PepsArgs A = args_of(m, m.theta, 0);
  hipLaunchKernelGGL(peps_env_kernel, dim3((unsigned)B), dim3(PT), 0, st, A, B, cfg, y0, y1,
                     dir, L.env);
  VMCX_LAUNCHED();
  return 0;
}

}  // namespace

int peps_forward(const vmcx_peps_model *m, int64_t B, const int64_t *cfg, void *ws,
                 int64_t ws_bytes, int64_t chunk, double *sign, double *logabs,
                 hipStream_t st) {
  if (check_peps(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && chunk >= 1, "bad sizes");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(cfg && sign && logabs && ws, "null pointer");
  const PepsWs L = peps_layout(*m, B, chunk, 0, (char *)ws);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes, "workspace too small");
  const int N = m->Lx * m->Ly;
  for (int64_t r0 = 0; r0 < B; r0 += chunk) {
    const int64_t rows = (B - r0 < chunk) ? B - r0 : chunk;
    const int rc = amp_rows(*m, L, rows, cfg + r0 * N, nullptr, sign + r0, logabs + r0, st);
    if (rc) return rc;
  }
  return 0;
}

int peps_sweep(const vmcx_peps_model *m, int64_t B, int n_edges, const int32_t *edges_host,
               double hopping_rate, int n_sweeps, uint64_t seed, int64_t sweep_offset,
               const double *u_choice, const uint8_t *bits, const double *u_acc, int64_t *cfg,
               double *sign, double *logabs, int64_t *n_accept, void *ws, int64_t ws_bytes,
               hipStream_t st) {
  if (check_peps(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && n_edges >= 0 && n_sweeps >= 0, "negative size");
  VMCX_CHECK_ARG((u_choice == nullptr) == (u_acc == nullptr) &&
                     (u_choice == nullptr) == (bits == nullptr),
                 "injected draws need u_choice, bits and u_acc together");
  if (B == 0 || n_sweeps == 0 || n_edges == 0) return 0;
  VMCX_CHECK_ARG(cfg && sign && logabs && ws && edges_host, "null pointer");
  const PepsWs L = peps_layout(*m, B, B, 0, (char *)ws);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes, "workspace too small (chunk must be >= B)");
  const int N = m->Lx * m->Ly;
  for (int e = 0; e < n_edges; ++e) {
    const int i = edges_host[2 * e], j = edges_host[2 * e + 1];
    VMCX_CHECK_ARG(i >= 0 && i < N && j >= 0 && j < N && i != j, "edge out of range");
  }
  const unsigned gw = (unsigned)((B + 255) / 256);
  const unsigned gc = (unsigned)((B * N + 255) / 256);
  // chi-truncated plain fPEPS: proposals resume from the walkers' cached
  // boundaries (bit-identical to the full contraction)
  const bool cached = use_bmps(*m) && m->hidden == 0 && !m->bmps_nocache;
  int32_t *acc = nullptr;
  if (cached) {
    const int rc = bmps_sweep_init(m->Lx, m->Ly, m->D, m->chi, m->dense_map, m->theta, B, cfg,
                                   L.bmc, st);
    if (rc) return rc;
    acc = bmps_sweep_accept_flags(m->Lx, m->Ly, m->D, m->chi, B, L.bmc);
  }
  for (int sw = 0; sw < n_sweeps; ++sw) {
    for (int e = 0; e < n_edges; ++e) {
      const int i = edges_host[2 * e], j = edges_host[2 * e + 1];
      const size_t di = ((size_t)sw * n_edges + e) * B;
      hipLaunchKernelGGL(nnbf_propose_kernel, dim3(gw), dim3(256), 0, st, B, N, i, j, e,
                         sweep_offset + sw, seed, 1.0 - hopping_rate,
                         u_choice ? u_choice + di : nullptr, bits ? bits + di : nullptr,
                         u_acc ? u_acc + di : nullptr, cfg, L.prop, L.moves, L.uacc);
      hipLaunchKernelGGL(peps_prop_cfg_kernel, dim3(gc), dim3(256), 0, st, B, N, i, j, cfg,
                         L.prop, L.pcfg);
      VMCX_LAUNCHED();
      const int ri = i / m->Ly, rj = j / m->Ly;
      int rc = cached ? bmps_sweep_amplitudes(m->Lx, m->Ly, m->D, m->chi, m->dense_map,
                                              m->theta, B, L.pcfg, L.prop, ri, rj, L.bmc,
                                              L.psign, L.plog, st)
                      : amp_rows(*m, L, B, L.pcfg, L.prop, L.psign, L.plog, st);
      if (rc) return rc;
      hipLaunchKernelGGL(peps_accept_kernel, dim3(gw), dim3(256), 0, st, B, N, i, j, cfg,
                         L.prop, L.uacc, L.psign, L.plog, sign, logabs, n_accept, acc);
      VMCX_LAUNCHED();
      if (cached) {
        rc = bmps_sweep_commit(m->Lx, m->Ly, m->D, m->chi, B, ri, rj, L.bmc, st);
        if (rc) return rc;
      }
    }
  }
  return 0;
}

int peps_local_energy(const vmcx_peps_model *m, int64_t B, int n_edges, const int32_t *edges,
                      double t, double U, const int64_t *cfg, const double *sign,
                      const double *logabs, void *ws, int64_t ws_bytes, int64_t chunk,
                      double *eloc, hipStream_t st) {
  if (check_peps(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && n_edges >= 0 && chunk >= 1, "bad sizes");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(cfg && sign && logabs && ws && eloc && (n_edges == 0 || edges),
                 "null pointer");
  const PepsWs L = peps_layout(*m, B, chunk, n_edges, (char *)ws);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes, "workspace too small");
  const int N = m->Lx * m->Ly;
  const unsigned gw = (unsigned)((B + 255) / 256);
  hipLaunchKernelGGL(conn_count_walker_kernel, dim3(gw), dim3(256), 0, st, B, N, n_edges,
                     edges, cfg, L.counts);
  hipLaunchKernelGGL(exclusive_scan_kernel, dim3(1), dim3(1024), 0, st, B, L.counts, L.offsets,
                     L.total);
  VMCX_LAUNCHED();
  const bool cached = use_bmps(*m) && m->hidden == 0 && !m->bmps_nocache;
  {
    const int rc = launch_conn_fill(B, N, n_edges, edges, t, cfg, L.offsets, L.etas, L.coeffs,
                                    st, cached ? L.cparent : nullptr, cached ? L.cmove : nullptr);
    if (rc) return rc;
  }
  int64_t T = 0;
  VMCX_HIP(hipMemcpyAsync(&T, L.total, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  VMCX_HIP(hipStreamSynchronize(st));
  if (cached) {
    // every connected configuration from its walker's cached boundaries
    int rc = bmps_sweep_init(m->Lx, m->Ly, m->D, m->chi, m->dense_map, m->theta, B, cfg, L.bmc,
                             st);
    if (rc) return rc;
    rc = bmps_conn_amplitudes(m->Lx, m->Ly, m->D, m->chi, m->dense_map, m->theta, B, T, L.etas,
                              L.cparent, L.cmove, L.bmc, L.csign, L.clog, st);
    if (rc) return rc;
  } else {
    for (int64_t r0 = 0; r0 < T; r0 += chunk) {
      const int64_t rows = (T - r0 < chunk) ? T - r0 : chunk;
      const int rc = amp_rows(*m, L, rows, L.etas + r0 * N, nullptr, L.csign + r0, L.clog + r0,
                              st);
      if (rc) return rc;
    }
  }
  hipLaunchKernelGGL(eloc_segment_kernel, dim3(gw), dim3(256), 0, st, B, N, U, cfg, L.offsets,
                     L.counts, L.coeffs, L.csign, L.clog, sign, logabs, eloc);
  VMCX_LAUNCHED();
  return 0;
}

// Metropolis sweeps with environment reuse: the same proposals, draws and
// Metropolis tests as peps_sweep, each proposal evaluated from the walker's
// cached environments (L_c0, R_{c1+1} for an edge spanning columns c0 .. c1)
// instead of a full contraction.  The cache is kept lazily: after an edge,
// L_y is stale for y > c0 and R_y for y <= c1 (those columns may have
// changed), and a stale environment is rebuilt incrementally from the
// nearest valid one when an edge needs it.  With the reference's edge order
// (row edges row by row, then column edges column by column) that costs
// about one column per edge on top of the proposal's own columns.
int peps_sweep_reuse(const vmcx_peps_model *m, int64_t B, int n_edges, const int32_t *edges_host,
                     double hopping_rate, int n_sweeps, uint64_t seed, int64_t sweep_offset,
                     const double *u_choice, const uint8_t *bits, const double *u_acc,
                     int64_t *cfg, double *sign, double *logabs, int64_t *n_accept, void *ws,
                     int64_t ws_bytes, hipStream_t st) {
  if (check_reuse(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && n_edges >= 0 && n_sweeps >= 0, "negative size");
  VMCX_CHECK_ARG((u_choice == nullptr) == (u_acc == nullptr) &&
                     (u_choice == nullptr) == (bits == nullptr),
                 "injected draws need u_choice, bits and u_acc together");
  if (B == 0 || n_sweeps == 0 || n_edges == 0) return 0;
  VMCX_CHECK_ARG(cfg && sign && logabs && ws && edges_host, "null pointer");
  const PepsWs L = peps_layout(*m, B, B, 0, (char *)ws, true);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes,
                 "workspace too small (vmcx_peps_reuse_workspace_bytes with chunk >= B)");
  const int N = m->Lx * m->Ly, Ly = m->Ly;
  for (int e = 0; e < n_edges; ++e) {
    const int i = edges_host[2 * e], j = edges_host[2 * e + 1];
    VMCX_CHECK_ARG(i >= 0 && i < N && j >= 0 && j < N && i != j, "edge out of range");
  }
  const unsigned gw = (unsigned)((B + 255) / 256);
  const unsigned gc = (unsigned)((B * N + 255) / 256);
  hipLaunchKernelGGL(peps_env_init_kernel, dim3(gw), dim3(256), 0, st, B, Ly, L.env);
  VMCX_LAUNCHED();
  const PepsArgs A = args_of(*m, m->theta, 0);
  int Lv = 0, Rv = Ly;   // L_y valid for y <= Lv, R_y for y >= Rv
  for (int sw = 0; sw < n_sweeps; ++sw) {
    for (int e = 0; e < n_edges; ++e) {
      const int i = edges_host[2 * e], j = edges_host[2 * e + 1];
      const int ci = i % Ly, cj = j % Ly;
      const int c0 = ci < cj ? ci : cj, c1 = ci < cj ? cj : ci;
      if (Lv < c0) {
        const int rc = env_sweep(*m, L, B, cfg, Lv, c0, 0, st);
        if (rc) return rc;
        Lv = c0;
      }
      if (Rv > c1 + 1) {
        const int rc = env_sweep(*m, L, B, cfg, c1 + 1, Rv, 1, st);
        if (rc) return rc;
        Rv = c1 + 1;
      }
      const size_t di = ((size_t)sw * n_edges + e) * B;
      hipLaunchKernelGGL(nnbf_propose_kernel, dim3(gw), dim3(256), 0, st, B, N, i, j, e,
                         sweep_offset + sw, seed, 1.0 - hopping_rate,
                         u_choice ? u_choice + di : nullptr, bits ? bits + di : nullptr,
                         u_acc ? u_acc + di : nullptr, cfg, L.prop, L.moves, L.uacc);
      hipLaunchKernelGGL(peps_prop_cfg_kernel, dim3(gc), dim3(256), 0, st, B, N, i, j, cfg,
                         L.prop, L.pcfg);
      hipLaunchKernelGGL(peps_amp_env_kernel, dim3((unsigned)B), dim3(PT), 0, st, A, B, L.pcfg,
                         nullptr, nullptr, c0, c1, L.prop, L.env, L.psign, L.plog);
      hipLaunchKernelGGL(peps_accept_kernel, dim3(gw), dim3(256), 0, st, B, N, i, j, cfg,
                         L.prop, L.uacc, L.psign, L.plog, sign, logabs, n_accept, nullptr);
      VMCX_LAUNCHED();
      if (Lv > c0) Lv = c0;
      if (Rv < c1 + 1) Rv = c1 + 1;
    }
  }
  return 0;
}

// E_loc with environment reuse: the walkers' environments once (2 Ly
// columns each), then every connected configuration (one hop) from its
// walker's environments, absorbing only the one or two columns the hop
// touches.
int peps_local_energy_reuse(const vmcx_peps_model *m, int64_t B, int n_edges,
                            const int32_t *edges, double t, double U, const int64_t *cfg,
                            const double *sign, const double *logabs, void *ws,
                            int64_t ws_bytes, int64_t chunk, double *eloc, hipStream_t st) {
  if (check_reuse(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && n_edges >= 0 && chunk >= 1, "bad sizes");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(cfg && sign && logabs && ws && eloc && (n_edges == 0 || edges),
                 "null pointer");
  const PepsWs L = peps_layout(*m, B, chunk, n_edges, (char *)ws, true);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes,
                 "workspace too small (vmcx_peps_reuse_workspace_bytes)");
  const int N = m->Lx * m->Ly, Ly = m->Ly;
  const unsigned gw = (unsigned)((B + 255) / 256);
  hipLaunchKernelGGL(conn_count_walker_kernel, dim3(gw), dim3(256), 0, st, B, N, n_edges,
                     edges, cfg, L.counts);
  hipLaunchKernelGGL(exclusive_scan_kernel, dim3(1), dim3(1024), 0, st, B, L.counts, L.offsets,
                     L.total);
  VMCX_LAUNCHED();
  {
    const int rc = launch_conn_fill(B, N, n_edges, edges, t, cfg, L.offsets, L.etas, L.coeffs,
                                    st, L.cparent, L.cmove);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(peps_env_init_kernel, dim3(gw), dim3(256), 0, st, B, Ly, L.env);
  VMCX_LAUNCHED();
  int rc = env_sweep(*m, L, B, cfg, 0, Ly, 0, st);
  if (rc) return rc;
  rc = env_sweep(*m, L, B, cfg, 0, Ly, 1, st);
  if (rc) return rc;
  int64_t T = 0;
  VMCX_HIP(hipMemcpyAsync(&T, L.total, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  VMCX_HIP(hipStreamSynchronize(st));
  const PepsArgs A = args_of(*m, m->theta, 0);
  for (int64_t r0 = 0; r0 < T; r0 += chunk) {
    const int64_t rows = (T - r0 < chunk) ? T - r0 : chunk;
    hipLaunchKernelGGL(peps_amp_env_kernel, dim3((unsigned)rows), dim3(PT), 0, st, A, rows,
                       L.etas + r0 * N, L.cparent + r0, L.cmove + r0, 0, 0, nullptr, L.env,
                       L.csign + r0, L.clog + r0);
    VMCX_LAUNCHED();
  }
  hipLaunchKernelGGL(eloc_segment_kernel, dim3(gw), dim3(256), 0, st, B, N, U, cfg, L.offsets,
                     L.counts, L.coeffs, L.csign, L.clog, sign, logabs, eloc);
  VMCX_LAUNCHED();
  return 0;
}

static int peps_logderiv_impl(const vmcx_peps_model *m, int64_t B, const int64_t *cfg,
                              void *ws, int64_t ws_bytes, int64_t chunk, void *O, int o_dtype,
                              int64_t ld, bool fact, double *sign, double *logabs,
                              hipStream_t st) {
  if (check_peps(m)) return 1;
  VMCX_CHECK_ARG(B >= 0 && chunk >= 1, "bad sizes");
  VMCX_CHECK_ARG(o_dtype == 0 || o_dtype == 1, "o_dtype must be 0 (f64) or 1 (f32)");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(cfg && O && sign && logabs && ws, "null pointer");
  const int64_t P = m->n_params, H = m->hidden;
  const int N = m->Lx * m->Ly;
  const bool mult = H > 0 && m->mode == 1;
  const int64_t np_total =
      fact ? (H > 0 ? P + 2 * N + 4 * H + (mult ? P : 0) : P)
           : (H > 0 ? P + H * 2 * N + H + H * H + H + P * H + P : P);
  VMCX_CHECK_ARG(ld >= np_total, fact ? "ld smaller than the factor width"
                                      : "ld smaller than the parameter count");
  const PepsWs L = peps_layout(*m, B, chunk, 0, (char *)ws);
  VMCX_CHECK_ARG((int64_t)L.bytes <= ws_bytes, "workspace too small");
  rocblas_handle h = blas_handle_for(st);
  VMCX_CHECK_ARG(h != nullptr, "rocblas handle");
  for (int64_t r0 = 0; r0 < B; r0 += chunk) {
    const int64_t rows = (B - r0 < chunk) ? B - r0 : chunk;
    const int64_t *c = cfg + r0 * N;
    PepsArgs A = args_of(*m, m->theta, 0);
    if (H > 0) {
      const int rc = mlp_theta(*m, L, rows, c, st);
      if (rc) return rc;
      A = args_of(*m, L.theta_eff, P);
    }
    // FP64 G (the upstream of the MLP backward); for plain fPEPS written
    // straight into an FP64 O
    const bool direct = (H == 0 && o_dtype == 0);
    double *G = direct ? (double *)O + r0 * ld : L.G;
    const int64_t ldg = direct ? ld : P;
    if (use_bmps(*m)) {
      // the reference's cheap gradient on the chi-truncated environments
      const int rc = bmps_logderiv(m->Lx, m->Ly, m->D, m->chi, A.dmap, A.theta, A.theta_stride,
                                   P, rows, c, G, ldg, sign + r0, logabs + r0, L.bm, L.bm_gpass,
                                   st);
      if (rc) return rc;
    } else {
      hipLaunchKernelGGL(peps_grad_kernel, dim3((unsigned)rows), dim3(PT), 0, st, A, rows, c,
                         G, ldg, sign + r0, logabs + r0);
      VMCX_LAUNCHED();
    }
    if (direct) continue;
    if (H == 0) {
      // FP32 O: plain copy of G
      hipLaunchKernelGGL(peps_copy_g_kernel, dim3((unsigned)rows), dim3(256), 0, st, rows, P, G,
                         (float *)O + r0 * ld, ld);
      VMCX_LAUNCHED();
      continue;
    }
    const int64_t nt = rows * P;
    if (mult) {
      // G -> G theta (upstream of the network), y -> G (1 + eta y) (theta block)
      const unsigned gt = (unsigned)((nt + 255) / 256 < 16384 ? (nt + 255) / 256 : 16384);
      hipLaunchKernelGGL(pnn_mul_grad_kernel, dim3(gt), dim3(256), 0, st, nt, P, m->eta,
                         m->theta, L.y, G);
      VMCX_LAUNCHED();
    }
    const double *Gth = mult ? L.y : G;
    // dh2 = (eta G) Wo ; dz2 = dh2 act'(z2) ; dh1 = dz2 W2 ; dz1 = dh1 act'(z1)
    const double etav = m->eta, one = 1.0, zero = 0.0;
    PEPS_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)H,
                            (rocblas_int)rows, (rocblas_int)P, &etav, m->Wo, (rocblas_int)H,
                            G, (rocblas_int)P, &zero, L.d2, (rocblas_int)H));
    const int64_t nh = rows * H;
    const unsigned gh = (unsigned)((nh + 255) / 256 < 8192 ? (nh + 255) / 256 : 8192);
    hipLaunchKernelGGL(pnn_dact_kernel, dim3(gh), dim3(256), 0, st, nh, m->act, L.z2, L.d2);
    VMCX_LAUNCHED();
    PEPS_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)H,
                            (rocblas_int)rows, (rocblas_int)H, &one, m->W2, (rocblas_int)H,
                            L.d2, (rocblas_int)H, &zero, L.d1, (rocblas_int)H));
    hipLaunchKernelGGL(pnn_dact_kernel, dim3(gh), dim3(256), 0, st, nh, m->act, L.z1, L.d1);
    VMCX_LAUNCHED();
    const size_t sm = (size_t)4 * H * 8;
    if (fact)
      hipLaunchKernelGGL((pnn_write_kernel<double, true>), dim3((unsigned)rows), dim3(256), sm,
                         st, rows, N, (int)H, P, m->act, m->eta, c, G, Gth, L.z1, L.z2, L.d1,
                         L.d2, (double *)O + r0 * ld, ld);
    else if (o_dtype == 0)
      hipLaunchKernelGGL(pnn_write_kernel<double>, dim3((unsigned)rows), dim3(256), sm, st, rows,
                         N, (int)H, P, m->act, m->eta, c, G, Gth, L.z1, L.z2, L.d1, L.d2,
                         (double *)O + r0 * ld, ld);
    else
      hipLaunchKernelGGL(pnn_write_kernel<float>, dim3((unsigned)rows), dim3(256), sm, st, rows,
                         N, (int)H, P, m->act, m->eta, c, G, Gth, L.z1, L.z2, L.d1, L.d2,
                         (float *)O + r0 * ld, ld);
    VMCX_LAUNCHED();
  }
  return 0;
}

int peps_logderiv(const vmcx_peps_model *m, int64_t B, const int64_t *cfg, void *ws,
                  int64_t ws_bytes, int64_t chunk, void *O, int o_dtype, int64_t ld,
                  double *sign, double *logabs, hipStream_t st) {
  return peps_logderiv_impl(m, B, cfg, ws, ws_bytes, chunk, O, o_dtype, ld, false, sign, logabs,
                            st);
}

int peps_logderiv_factors(const vmcx_peps_model *m, int64_t B, const int64_t *cfg, void *ws,
                          int64_t ws_bytes, int64_t chunk, double *F, int64_t ld,
                          double *sign, double *logabs, hipStream_t st) {
  return peps_logderiv_impl(m, B, cfg, ws, ws_bytes, chunk, F, 0, ld, true, sign, logabs, st);
}

}  // namespace vmcx
