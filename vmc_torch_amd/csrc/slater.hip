// Slater-determinant walker kernels for gfx950 (CDNA4).  The walker state
// A^-1 is stored in FP64 (the plain form) or FP32 (the precision form, the
// `T = float` instances); the exact refresh always runs in FP64.
//
// Layout: a walker owns a group of NFP lanes (NFP = 8/16/32/64, so 64/NFP
// walkers share one wave64).  Lane k of the group holds column k of A^-1 in
// NFP registers (A = M[occupied orbitals, :], row k = slot k).  With this
// layout:
//   * the ratio row v = m_b^T A^-1 is one lane-local dot product per lane
//     against a broadcast row of M held in LDS;
//   * a Sherman-Morrison update is one broadcast column (LDS) and NFP FMAs
//     per lane;
//   * the log-derivative row of an occupied orbital is a contiguous column
//     of A^-1 in HBM (column k stored at ainv + (b*NFP + k)*NFP).
#include "gj.h"
#include "kernels.h"

namespace vmcx {

template <int NFP>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = NFP / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int NFP>
__device__ __forceinline__ uint64_t group_or(uint64_t v) {
#pragma unroll
  for (int o = NFP / 2; o > 0; o >>= 1) {
    const uint32_t lo = __shfl_xor((uint32_t)v, o, 64);
    const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), o, 64);
    v |= ((uint64_t)hi << 32) | lo;
  }
  return v;
}

// Up / down occupation masks of walker w, reduced over the group's lanes.
template <int NFP>
__device__ __forceinline__ void load_masks(const int64_t *cfg, int n_sites,
                                           int col, uint64_t &up,
                                           uint64_t &dn) {
  uint64_t u = 0, d = 0;
  for (int s = col; s < n_sites; s += NFP) {
    const int64_t x = cfg[s];
    if (x & 2) u |= 1ull << s;
    if (x & 1) d |= 1ull << s;
  }
  up = group_or<NFP>(u);
  dn = group_or<NFP>(d);
}

// ======================================================== exact refresh
// Gauss-Jordan with partial pivoting on B = A^T held row-per-lane (lane k =
// row k of B = column k of A, register j = row j of A).  Pivot rows are
// lanes; the pivot column is always register 0 because the register file is
// rotated by one after each step (NFP rotations restore the order).  The
// in-place scheme leaves R[q_p][c] = B^-1[p][q_c] = A^-1[q_c][p] in lane
// q_p, register c, which is scattered to ainv[p][q_c] on the way out.
// INV = false (no ainv requested): the determinant alone by LU (lu_logdet).
template <int NFP, bool INV, typename T>
__global__ void __launch_bounds__(512)
    refresh_kernel(int64_t B, int n_sites, int n_f, const double *__restrict__ M,
                   const int64_t *__restrict__ configs, T *ainv,
                   uint8_t *slot, double *sign, double *logabs,
                   int32_t *since_refresh, int only_if_since) {
  constexpr int WPW = 64 / NFP;
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = lane / NFP, col = lane % NFP;
  const int wl = wave * WPW + grp;
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x / 64) * WPW + wl;
  if (w >= B) return;
  if (only_if_since >= 0 && since_refresh != nullptr &&
      since_refresh[w] < only_if_since)
    return;

  double *prow = reinterpret_cast<double *>(smem) + wl * NFP;
  uint8_t *qbuf = reinterpret_cast<uint8_t *>(smem) +
                  (size_t)(blockDim.x / 64) * WPW * NFP * sizeof(double) +
                  wl * NFP;

  uint64_t up, dn;
  load_masks<NFP>(configs + w * n_sites, n_sites, col, up, dn);
  const int nup = __popcll(up);
  const bool count_ok = (nup + __popcll(dn)) == n_f;

  // Row `col` of B: a[j] = A[j][col] = M[orb_j][col] (orb_j ascending).
  double a[NFP];
  {
    uint64_t mu = up, md = dn;
#pragma unroll
    for (int j = 0; j < NFP; ++j) {
      int orb = -1;
      if (mu) {
        orb = __ffsll((long long)mu) - 1;
        mu &= mu - 1;
      } else if (md) {
        orb = n_sites + __ffsll((long long)md) - 1;
        md &= md - 1;
      }
      if (j < n_f && col < n_f)
        a[j] = (orb >= 0) ? M[(int64_t)orb * n_f + col] : 0.0;
      else
        a[j] = (j == col) ? 1.0 : 0.0;
    }
  }

  bool singular = !count_ok;
  const bool nonfinite = group_nonfinite<NFP>(a, grp * NFP);
  const GJResult r = INV ? gj_invert<NFP>(a, col, grp * NFP, prow, qbuf)
                         : lu_logdet<NFP>(a, col, grp * NFP, prow);
  singular = singular || r.singular;
  const double logdet = r.logdet;
  const int parity = r.parity;
  const int my_step = r.my_step;
  if (col == 0) {
    if (nonfinite) {
      sign[w] = __builtin_nan("");
      logabs[w] = __builtin_nan("");
    } else if (singular) {
      sign[w] = 0.0;
      logabs[w] = count_ok ? log(1e-45) : __builtin_nan("");
    } else {
      sign[w] = (parity & 1) ? -1.0 : 1.0;
      logabs[w] = fmax(logdet, log(1e-45));
    }
    if (since_refresh) since_refresh[w] = 0;
  }
  if (INV && my_step >= 0) {
    // lane q_p holds column p of A^-1 with register c = row q_c
    T *dst = ainv + ((size_t)w * NFP + my_step) * NFP;
#pragma unroll
    for (int c = 0; c < NFP; ++c)
      dst[qbuf[c]] = nonfinite ? T(__builtin_nan("")) : (singular ? T(0) : (T)a[c]);
  }
  if (slot != nullptr) {
    const int n_orb = 2 * n_sites;
    for (int o = col; o < n_orb; o += NFP) {
      uint8_t s = 0xFF;
      if (o < n_sites) {
        if ((up >> o) & 1ull) s = (uint8_t)__popcll(up & ((1ull << o) - 1ull));
      } else {
        const int si = o - n_sites;
        if ((dn >> si) & 1ull)
          s = (uint8_t)(nup + __popcll(dn & ((1ull << si) - 1ull)));
      }
      slot[(size_t)w * n_orb + o] = s;
    }
  }
}

// ================================================================ sweep
// One wave group per walker; M (2N x NFP, zero padded) staged in LDS.
template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
  using type = double2;
};
template <>
struct Vec2<float> {
  using type = float2;
};

template <int NFP, typename T>
__device__ __forceinline__ T dot_row(const T *__restrict__ mrow, const T (&a)[NFP]) {
  using T2 = typename Vec2<T>::type;
  const T2 *m2 = reinterpret_cast<const T2 *>(mrow);
  T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
  for (int j = 0; j < NFP / 2; j += 2) {
    const T2 x = m2[j];
    const T2 y = m2[j + 1];
    s0 = fma(x.x, a[2 * j], s0);
    s1 = fma(x.y, a[2 * j + 1], s1);
    s2 = fma(y.x, a[2 * j + 2], s2);
    s3 = fma(y.y, a[2 * j + 3], s3);
  }
  return (s0 + s1) + (s2 + s3);
}


// per-walker LDS: two broadcast columns of A^-1 (2 NFP values), the site
// states (64 bytes) and the slots (128 bytes)
template <int NFP, typename T>
constexpr int walker_lds_bytes() {
  return 2 * NFP * (int)sizeof(T) + 64 + 128;
}

// the sweep's M (LDS copy) and A^-1 of the storage type T
template <typename T>
__device__ __forceinline__ const T *sweep_m(const SweepArgs &g);
template <>
__device__ __forceinline__ const double *sweep_m<double>(const SweepArgs &g) {
  return g.M;
}
template <>
__device__ __forceinline__ const float *sweep_m<float>(const SweepArgs &g) {
  return g.M32;
}
template <typename T>
__device__ __forceinline__ T *sweep_ainv(const SweepArgs &g);
template <>
__device__ __forceinline__ double *sweep_ainv<double>(const SweepArgs &g) {
  return g.ainv;
}
template <>
__device__ __forceinline__ float *sweep_ainv<float>(const SweepArgs &g) {
  return g.ainv32;
}

// T = double: the plain FP64 form.  T = float: ratios, determinant ratios
// and Sherman-Morrison updates in FP32 (half the registers and LDS per
// walker); the acceptance test, the sign and ln|psi| in FP64.
template <int NFP, typename T>
__global__ void __launch_bounds__(512) sweep_kernel(SweepArgs g) {
  constexpr int WPW = 64 / NFP;
  using T2 = typename Vec2<T>::type;
  extern __shared__ __align__(16) char smem[];
  const int N = g.n_sites, n_orb = 2 * N;
  T *Ms = reinterpret_cast<T *>(smem);
  const T *Mg = sweep_m<T>(g);
  T *ainv = sweep_ainv<T>(g);
  for (int idx = threadIdx.x; idx < n_orb * NFP; idx += blockDim.x) {
    const int r = idx / NFP, c = idx % NFP;
    Ms[idx] = (c < g.n_f) ? Mg[(int64_t)r * g.n_f + c] : T(0);
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = lane / NFP, col = lane % NFP;
  const int wl = wave * WPW + grp;
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x / 64) * WPW + wl;
  if (w >= g.B) return;
  const int gbase = grp * NFP;

  char *wbase = smem + (size_t)n_orb * NFP * sizeof(T) +
                (size_t)wl * walker_lds_bytes<NFP, T>();
  T *colbuf = reinterpret_cast<T *>(wbase);  // [2][NFP]
  uint8_t *st = reinterpret_cast<uint8_t *>(wbase + 2 * NFP * sizeof(T));
  uint8_t *sl = st + 64;

  T a[NFP];
  {
    const T2 *src = reinterpret_cast<const T2 *>(ainv + ((size_t)w * NFP + col) * NFP);
#pragma unroll
    for (int j = 0; j < NFP / 2; ++j) {
      const T2 v = src[j];
      a[2 * j] = v.x;
      a[2 * j + 1] = v.y;
    }
  }
  int64_t *cfg = g.configs + w * N;
  for (int s = col; s < N; s += NFP) st[s] = (uint8_t)cfg[s];
  for (int o = col; o < n_orb; o += NFP) sl[o] = g.slot[(size_t)w * n_orb + o];
  uint64_t up, dn;
  load_masks<NFP>(cfg, N, col, up, dn);
  wave_lds_fence();

  double sgn = g.sign[w], la = g.logabs[w];
  int since = g.since_refresh[w];
  int64_t nacc = 0;
  const uint32_t k0 = (uint32_t)g.seed, k1 = (uint32_t)(g.seed >> 32);
  const int64_t off0 = g.sweep_offset + (g.offset_dev ? g.offset_dev[0] : 0);

#pragma unroll 1
  for (int sw = 0; sw < g.n_sweeps; ++sw) {
#pragma unroll 1
    for (int e = 0; e < g.n_edges; ++e) {
      const int i = g.edges[2 * e], j = g.edges[2 * e + 1];
      if ((unsigned)i >= (unsigned)N || (unsigned)j >= (unsigned)N || i == j) continue;
      const int si = st[i], sj = st[j];
      if (si == sj) continue;
      double uc, ua;
      int b0, b2;
      if (g.u_choice == nullptr) {
        const u32x4 o = philox4x32_10((uint32_t)e, (uint32_t)(off0 + sw), (uint32_t)w, 0u,
                                      k0, k1);
        uc = u53(o.x, o.y);
        ua = u53(o.z, o.w);
        b0 = o.y & 1u;
        b2 = (o.y >> 1) & 1u;
      } else {
        const size_t di = ((size_t)sw * g.n_edges + e) * g.B + w;
        uc = g.u_choice[di];
        ua = g.u_acc[di];
        const int bb = g.bits[di];
        b0 = bb & 1;
        b2 = (bb >> 1) & 1;
      }
      // proposal (GPU/sampler.py:47-137)
      int ni, nj;
      const int nmi = (si == 0) ? 0 : ((si == 3) ? 2 : 1);
      const int nmj = (sj == 0) ? 0 : ((sj == 3) ? 2 : 1);
      const int dnum = nmi > nmj ? nmi - nmj : nmj - nmi;
      const bool is_ex = uc < g.keep_rate;
      if (is_ex || dnum == 1) {
        ni = sj;
        nj = si;
      } else if (dnum == 0) {
        ni = b0 ? 3 : 0;
        nj = b0 ? 0 : 3;
      } else {
        ni = b2 ? 2 : 1;
        nj = b2 ? 1 : 2;
      }
      // electrons that move: (orbital from -> orbital to)
      int af[2], bt[2], spin[2], nm = 0;
      if ((si ^ ni) & 2) {  // up electron moves
        const int from = (si & 2) ? i : j, to = (si & 2) ? j : i;
        af[nm] = from;
        bt[nm] = to;
        spin[nm] = 0;
        ++nm;
      }
      if ((si ^ ni) & 1) {  // down electron moves
        const int from = (si & 1) ? i : j, to = (si & 1) ? j : i;
        af[nm] = N + from;
        bt[nm] = N + to;
        spin[nm] = 1;
        ++nm;
      }
      const int kk1 = sl[af[0]];
      const T v1 = dot_row<NFP, T>(Ms + bt[0] * NFP, a);
      const T r11 = __shfl(v1, gbase + kk1, 64);
      T det, v2 = 0, r12 = 0, r21 = 0, r22 = 0;
      int kk2 = 0;
      if (nm == 2) {
        kk2 = sl[af[1]];
        v2 = dot_row<NFP, T>(Ms + bt[1] * NFP, a);
        r12 = __shfl(v1, gbase + kk2, 64);
        r21 = __shfl(v2, gbase + kk1, 64);
        r22 = __shfl(v2, gbase + kk2, 64);
        det = r11 * r22 - r12 * r21;
      } else {
        det = r11;
      }
      const double detd = (double)det;
      if (!(ua < detd * detd)) continue;

      // ---- accepted: Sherman-Morrison / Woodbury update of A^-1
      if (col == kk1) {
#pragma unroll
        for (int jj = 0; jj < NFP; ++jj) colbuf[jj] = a[jj];
      }
      if (nm == 2 && col == kk2) {
#pragma unroll
        for (int jj = 0; jj < NFP; ++jj) colbuf[NFP + jj] = a[jj];
      }
      wave_lds_fence();
      const T rdet = T(1) / det;
      if (nm == 1) {
        const T c1 = (v1 - (col == kk1 ? T(1) : T(0))) * rdet;
#pragma unroll
        for (int jj = 0; jj < NFP; ++jj) a[jj] = fma(-colbuf[jj], c1, a[jj]);
      } else {
        const T w1 = v1 - (col == kk1 ? T(1) : T(0));
        const T w2 = v2 - (col == kk2 ? T(1) : T(0));
        const T c1 = (r22 * w1 - r12 * w2) * rdet;
        const T c2 = (r11 * w2 - r21 * w1) * rdet;
#pragma unroll
        for (int jj = 0; jj < NFP; ++jj)
          a[jj] = fma(-colbuf[jj], c1, fma(-colbuf[NFP + jj], c2, a[jj]));
      }
      // netket row-order sign: same-spin occupied orbitals passed over
      int cnt = 0;
      for (int m = 0; m < nm; ++m) {
        const int sa = af[m] - spin[m] * N, sb = bt[m] - spin[m] * N;
        cnt += __popcll((spin[m] ? dn : up) & between_mask(sa, sb));
      }
      sgn *= ((detd < 0.0) != (cnt & 1)) ? -1.0 : 1.0;
      la += log(fabs(detd));
      // occupation and slot bookkeeping
      for (int m = 0; m < nm; ++m) {
        const int sa = af[m] - spin[m] * N, sb = bt[m] - spin[m] * N;
        uint64_t &mk = spin[m] ? dn : up;
        mk = (mk & ~(1ull << sa)) | (1ull << sb);
      }
      wave_lds_fence();
      if (col == 0) {
        st[i] = (uint8_t)ni;
        st[j] = (uint8_t)nj;
        sl[bt[0]] = (uint8_t)kk1;
        sl[af[0]] = 0xFF;
        if (nm == 2) {
          sl[bt[1]] = (uint8_t)kk2;
          sl[af[1]] = 0xFF;
        }
      }
      wave_lds_fence();
      ++since;
      ++nacc;
    }
  }

  {
    T2 *dst = reinterpret_cast<T2 *>(ainv + ((size_t)w * NFP + col) * NFP);
#pragma unroll
    for (int jj = 0; jj < NFP / 2; ++jj) {
      T2 v;
      v.x = a[2 * jj];
      v.y = a[2 * jj + 1];
      dst[jj] = v;
    }
  }
  for (int s = col; s < N; s += NFP) cfg[s] = st[s];
  for (int o = col; o < n_orb; o += NFP) g.slot[(size_t)w * n_orb + o] = sl[o];
  if (col == 0) {
    g.sign[w] = sgn;
    g.logabs[w] = la;
    g.since_refresh[w] = since;
    if (g.n_accept) g.n_accept[w] += nacc;
  }
}

// ========================================================= local energy
// T: the storage type of M and A^-1 (the ratios are formed in T, the sum
// over connected configurations in FP64)
template <int NFP, typename T>
__global__ void __launch_bounds__(512)
    local_energy_kernel(int64_t B, int n_sites, int n_f,
                        const T *__restrict__ M, int n_edges,
                        const int32_t *__restrict__ edges, double t, double U,
                        const int64_t *__restrict__ configs,
                        const T *__restrict__ ainv,
                        const uint8_t *__restrict__ slot, double *eloc) {
  constexpr int WPW = 64 / NFP;
  using T2 = typename Vec2<T>::type;
  extern __shared__ __align__(16) char smem[];
  const int N = n_sites, n_orb = 2 * N;
  T *Ms = reinterpret_cast<T *>(smem);
  uint64_t *nbr = reinterpret_cast<uint64_t *>(smem + (size_t)n_orb * NFP * sizeof(T));
  for (int idx = threadIdx.x; idx < n_orb * NFP; idx += blockDim.x) {
    const int r = idx / NFP, c = idx % NFP;
    Ms[idx] = (c < n_f) ? M[(int64_t)r * n_f + c] : T(0);
  }
  for (int s = threadIdx.x; s < N; s += blockDim.x) {
    uint64_t m = 0;
    for (int e = 0; e < n_edges; ++e) {
      const int i = edges[2 * e], j = edges[2 * e + 1];
      if ((unsigned)i >= (unsigned)N || (unsigned)j >= (unsigned)N || i == j) continue;
      if (i == s) m |= 1ull << j;
      if (j == s) m |= 1ull << i;
    }
    nbr[s] = m;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = lane / NFP, col = lane % NFP;
  const int wl = wave * WPW + grp;
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x / 64) * WPW + wl;
  if (w >= B) return;
  int *orbs = reinterpret_cast<int *>(nbr + 64) + wl * NFP;

  T a[NFP];
  {
    const T2 *src = reinterpret_cast<const T2 *>(ainv + ((size_t)w * NFP + col) * NFP);
#pragma unroll
    for (int j = 0; j < NFP / 2; ++j) {
      const T2 v = src[j];
      a[2 * j] = v.x;
      a[2 * j + 1] = v.y;
    }
  }
  orbs[col] = -1;
  wave_lds_fence();
  for (int o = col; o < n_orb; o += NFP) {
    const uint8_t s = slot[(size_t)w * n_orb + o];
    if (s != 0xFF && s < NFP) orbs[s] = o;
  }
  uint64_t up, dn;
  load_masks<NFP>(configs + w * N, N, col, up, dn);
  wave_lds_fence();
  const int my_orb = orbs[col];
  const int my_spin = (my_orb >= N) ? 1 : 0;
  const int my_site = my_orb - my_spin * N;

  double acc = 0.0;
  for (int b = 0; b < n_orb; ++b) {
    const int bspin = (b >= N) ? 1 : 0, bsite = b - bspin * N;
    const uint64_t occ = bspin ? dn : up;
    if ((occ >> bsite) & 1ull) continue;          // target must be empty
    if ((nbr[bsite] & occ) == 0ull) continue;     // needs an occupied neighbour
    const double v = (double)dot_row<NFP, T>(Ms + b * NFP, a);
    if (my_orb >= 0 && my_spin == bspin && ((nbr[bsite] >> my_site) & 1ull)) {
      const int lo = my_site < bsite ? my_site : bsite;
      const int hi = my_site < bsite ? bsite : my_site;
      const uint64_t mid = between_mask(lo, hi);
      const int par = bspin ? __popcll(up & (mid | (1ull << lo)))
                            : __popcll(dn & (mid | (1ull << hi)));
      acc += ((par & 1) ? t : -t) * v;
    }
  }
  const double tot = group_sum<NFP>(acc);
  if (col == 0) eloc[w] = tot + U * (double)__popcll(up & dn);
}

// ======================================================= log-derivative
template <typename T, typename TA>
__global__ void __launch_bounds__(256)
    logderiv_kernel(int64_t B, int n_sites, int n_f, int nfp,
                    const TA *__restrict__ ainv,
                    const uint8_t *__restrict__ slot, T *O, int64_t ld) {
  __shared__ uint8_t sl[128];
  const int64_t w = blockIdx.x;
  const int n_orb = 2 * n_sites;
  for (int o = threadIdx.x; o < n_orb; o += blockDim.x)
    sl[o] = slot[(size_t)w * n_orb + o];
  __syncthreads();
  const int total = n_orb * n_f;
  T *row = O + w * ld;
  const TA *base = ainv + (size_t)w * nfp * nfp;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int i = idx / n_f, j = idx - i * n_f;
    const uint8_t s = sl[i];
    row[idx] = (s == 0xFF) ? T(0) : (T)base[(size_t)s * nfp + j];
  }
}

// ============================================================ launchers
// key of the launch-shape table: the padded width and the storage precision
template <int NFP, typename T>
constexpr int tune_key() {
  return 2 * NFP + (sizeof(T) == 4 ? 1 : 0);
}

template <int NFP, typename T>
static int launch_refresh(int64_t B, int n_sites, int n_f, const double *M,
                          const int64_t *configs, T *ainv, uint8_t *slot,
                          double *sign, double *logabs, int32_t *since,
                          int only_if_since, hipStream_t st) {
  constexpr int WPW = 64 / NFP;
  const int waves = tuned_waves(TUNE_SLATER_REFRESH, tune_key<NFP, T>(), 4);
  const int64_t per_block = waves * WPW;
  const int64_t nblk = (B + per_block - 1) / per_block;
  const size_t lds = (size_t)per_block * NFP * (sizeof(double) + 1);
  if (ainv)
    hipLaunchKernelGGL((refresh_kernel<NFP, true, T>), dim3((unsigned)nblk), dim3(64 * waves),
                       lds, st, B, n_sites, n_f, M, configs, ainv, slot, sign, logabs, since,
                       only_if_since);
  else
    hipLaunchKernelGGL((refresh_kernel<NFP, false, T>), dim3((unsigned)nblk), dim3(64 * waves),
                       lds, st, B, n_sites, n_f, M, configs, ainv, slot, sign, logabs, since,
                       only_if_since);
  VMCX_LAUNCHED();
  return 0;
}

template <typename T>
static int refresh_t(int64_t B, int n_sites, int n_f, const double *M, const int64_t *configs,
                     T *ainv, uint8_t *slot, double *sign, double *logabs, int32_t *since,
                     int only_if_since, hipStream_t st) {
  VMCX_CHECK_ARG(B >= 0, "B < 0");
  VMCX_CHECK_ARG(n_sites >= 1 && n_sites <= 64, "n_sites must be in [1, 64]");
  VMCX_CHECK_ARG(n_f >= 1 && n_f <= 64, "n_f must be in [1, 64]");
  VMCX_CHECK_ARG(n_f <= 2 * n_sites, "n_f > 2*n_sites");
  VMCX_CHECK_ARG(M && configs && sign && logabs, "null pointer");
  VMCX_CHECK_ARG((ainv == nullptr) == (slot == nullptr),
                 "ainv and slot must both be given or both be NULL");
  if (B == 0) return 0;
  switch (nfp_of(n_f)) {
    case 8: return launch_refresh<8, T>(B, n_sites, n_f, M, configs, ainv, slot, sign, logabs, since, only_if_since, st);
    case 16: return launch_refresh<16, T>(B, n_sites, n_f, M, configs, ainv, slot, sign, logabs, since, only_if_since, st);
    case 32: return launch_refresh<32, T>(B, n_sites, n_f, M, configs, ainv, slot, sign, logabs, since, only_if_since, st);
    default: return launch_refresh<64, T>(B, n_sites, n_f, M, configs, ainv, slot, sign, logabs, since, only_if_since, st);
  }
}

int slater_refresh(int64_t B, int n_sites, int n_f, const double *M,
                   const int64_t *configs, double *ainv, uint8_t *slot,
                   double *sign, double *logabs, int32_t *since,
                   int only_if_since, hipStream_t st) {
  return refresh_t<double>(B, n_sites, n_f, M, configs, ainv, slot, sign, logabs, since,
                           only_if_since, st);
}

int slater_refresh_f32(int64_t B, int n_sites, int n_f, const double *M,
                       const int64_t *configs, float *ainv, uint8_t *slot, double *sign,
                       double *logabs, int32_t *since, int only_if_since, hipStream_t st) {
  return refresh_t<float>(B, n_sites, n_f, M, configs, ainv, slot, sign, logabs, since,
                          only_if_since, st);
}

template <int NFP, typename T>
static int launch_sweep(const SweepArgs &a, hipStream_t st) {
  constexpr int WPW = 64 / NFP;
  const int waves = tuned_waves(TUNE_SLATER_SWEEP, tune_key<NFP, T>(), 4);
  const int64_t per_block = waves * WPW;
  const int64_t nblk = (a.B + per_block - 1) / per_block;
  const size_t lds = (size_t)2 * a.n_sites * NFP * sizeof(T) +
                     (size_t)per_block * walker_lds_bytes<NFP, T>();
  VMCX_CHECK_ARG(lds <= max_lds_per_block(), "LDS request exceeds the device limit");
  hipLaunchKernelGGL((sweep_kernel<NFP, T>), dim3((unsigned)nblk), dim3(64 * waves), lds, st,
                     a);
  VMCX_LAUNCHED();
  return 0;
}

// T = double: a.M / a.ainv; T = float: a.M32 / a.ainv32 for the sweep and
// a.M (FP64) for the exact refresh between sweeps
template <typename T>
static int sweep_t(const SweepArgs &a, int refresh_every, hipStream_t st) {
  VMCX_CHECK_ARG(a.B >= 0 && a.n_sweeps >= 0 && a.n_edges >= 0, "negative size");
  VMCX_CHECK_ARG(a.n_sites >= 1 && a.n_sites <= 64, "n_sites must be in [1, 64]");
  VMCX_CHECK_ARG(a.n_f >= 1 && a.n_f <= 64 && a.n_f <= 2 * a.n_sites, "bad n_f");
  const bool f32 = sizeof(T) == 4;
  VMCX_CHECK_ARG(a.M && a.configs && (f32 ? (a.ainv32 && a.M32) : a.ainv != nullptr) &&
                     a.slot && a.sign && a.logabs && a.since_refresh,
                 "null state pointer");
  VMCX_CHECK_ARG(a.n_edges == 0 || a.edges, "null edges");
  VMCX_CHECK_ARG((a.u_choice == nullptr) == (a.u_acc == nullptr) &&
                     (a.u_choice == nullptr) == (a.bits == nullptr),
                 "injected draws need u_choice, bits and u_acc together");
  if (a.B == 0 || a.n_sweeps == 0) return 0;
  const int nfp = nfp_of(a.n_f);
  SweepArgs one = a;
  // one launch per sweep so that the exact refresh runs between sweeps
  const int per_launch = (refresh_every > 0) ? 1 : a.n_sweeps;
  for (int s0 = 0; s0 < a.n_sweeps; s0 += per_launch) {
    one.n_sweeps = (a.n_sweeps - s0 < per_launch) ? a.n_sweeps - s0 : per_launch;
    one.sweep_offset = a.sweep_offset + s0;
    if (a.u_choice) {
      const size_t off = (size_t)s0 * a.n_edges * a.B;
      one.u_choice = a.u_choice + off;
      one.u_acc = a.u_acc + off;
      one.bits = a.bits + off;
    }
    int rc;
    switch (nfp) {
      case 8: rc = launch_sweep<8, T>(one, st); break;
      case 16: rc = launch_sweep<16, T>(one, st); break;
      case 32: rc = launch_sweep<32, T>(one, st); break;
      default: rc = launch_sweep<64, T>(one, st); break;
    }
    if (rc) return rc;
    if (refresh_every > 0) {
      T *ainv = f32 ? (T *)a.ainv32 : (T *)a.ainv;
      rc = refresh_t<T>(a.B, a.n_sites, a.n_f, a.M, a.configs, ainv, a.slot, a.sign,
                        a.logabs, a.since_refresh, refresh_every, st);
      if (rc) return rc;
    }
  }
  return 0;
}

int slater_sweep(const SweepArgs &a, int refresh_every, hipStream_t st) {
  return sweep_t<double>(a, refresh_every, st);
}

int slater_sweep_f32(const SweepArgs &a, int refresh_every, hipStream_t st) {
  return sweep_t<float>(a, refresh_every, st);
}

template <int NFP, typename T>
static int launch_eloc(int64_t B, int n_sites, int n_f, const T *M, int n_edges,
                       const int32_t *edges, double t, double U, const int64_t *configs,
                       const T *ainv, const uint8_t *slot, double *eloc, hipStream_t st) {
  constexpr int WPW = 64 / NFP;
  const int waves = tuned_waves(TUNE_SLATER_ELOC, tune_key<NFP, T>(), 4);
  const int64_t per_block = waves * WPW;
  const int64_t nblk = (B + per_block - 1) / per_block;
  const size_t lds = (size_t)2 * n_sites * NFP * sizeof(T) +
                     64 * sizeof(uint64_t) + (size_t)per_block * NFP * sizeof(int);
  VMCX_CHECK_ARG(lds <= max_lds_per_block(), "LDS request exceeds the device limit");
  hipLaunchKernelGGL((local_energy_kernel<NFP, T>), dim3((unsigned)nblk), dim3(64 * waves),
                     lds, st, B, n_sites, n_f, M, n_edges, edges, t, U, configs, ainv, slot,
                     eloc);
  VMCX_LAUNCHED();
  return 0;
}

template <typename T>
static int eloc_t(int64_t B, int n_sites, int n_f, const T *M, int n_edges,
                  const int32_t *edges, double t, double U, const int64_t *configs,
                  const T *ainv, const uint8_t *slot, double *eloc, hipStream_t st) {
  VMCX_CHECK_ARG(B >= 0 && n_edges >= 0, "negative size");
  VMCX_CHECK_ARG(n_sites >= 1 && n_sites <= 64, "n_sites must be in [1, 64]");
  VMCX_CHECK_ARG(n_f >= 1 && n_f <= 64 && n_f <= 2 * n_sites, "bad n_f");
  VMCX_CHECK_ARG(M && configs && ainv && slot && eloc, "null pointer");
  VMCX_CHECK_ARG(n_edges == 0 || edges, "null edges");
  if (B == 0) return 0;
  switch (nfp_of(n_f)) {
    case 8: return launch_eloc<8, T>(B, n_sites, n_f, M, n_edges, edges, t, U, configs, ainv, slot, eloc, st);
    case 16: return launch_eloc<16, T>(B, n_sites, n_f, M, n_edges, edges, t, U, configs, ainv, slot, eloc, st);
    case 32: return launch_eloc<32, T>(B, n_sites, n_f, M, n_edges, edges, t, U, configs, ainv, slot, eloc, st);
    default: return launch_eloc<64, T>(B, n_sites, n_f, M, n_edges, edges, t, U, configs, ainv, slot, eloc, st);
  }
}

int slater_local_energy(int64_t B, int n_sites, int n_f, const double *M,
                        int n_edges, const int32_t *edges, double t, double U,
                        const int64_t *configs, const double *ainv,
                        const uint8_t *slot, double *eloc, hipStream_t st) {
  return eloc_t<double>(B, n_sites, n_f, M, n_edges, edges, t, U, configs, ainv, slot, eloc,
                        st);
}

int slater_local_energy_f32(int64_t B, int n_sites, int n_f, const float *M, int n_edges,
                            const int32_t *edges, double t, double U, const int64_t *configs,
                            const float *ainv, const uint8_t *slot, double *eloc,
                            hipStream_t st) {
  return eloc_t<float>(B, n_sites, n_f, M, n_edges, edges, t, U, configs, ainv, slot, eloc,
                       st);
}

template <typename TA>
static int logderiv_t(int64_t B, int n_sites, int n_f, const TA *ainv, const uint8_t *slot,
                      void *O, int o_dtype, int64_t ld, hipStream_t st) {
  VMCX_CHECK_ARG(B >= 0, "B < 0");
  VMCX_CHECK_ARG(n_sites >= 1 && n_sites <= 64, "n_sites must be in [1, 64]");
  VMCX_CHECK_ARG(n_f >= 1 && n_f <= 64 && n_f <= 2 * n_sites, "bad n_f");
  VMCX_CHECK_ARG(ld >= (int64_t)2 * n_sites * n_f, "ld < 2*n_sites*n_f");
  VMCX_CHECK_ARG(ainv && slot && O, "null pointer");
  VMCX_CHECK_ARG(o_dtype == 0 || o_dtype == 1, "o_dtype must be 0 (f64) or 1 (f32)");
  VMCX_CHECK_ARG(B <= 0x7fffffff, "B too large for one grid");
  if (B == 0) return 0;
  const int nfp = nfp_of(n_f);
  if (o_dtype == 0)
    hipLaunchKernelGGL((logderiv_kernel<double, TA>), dim3((unsigned)B), dim3(256), 0, st, B,
                       n_sites, n_f, nfp, ainv, slot, (double *)O, ld);
  else
    hipLaunchKernelGGL((logderiv_kernel<float, TA>), dim3((unsigned)B), dim3(256), 0, st, B,
                       n_sites, n_f, nfp, ainv, slot, (float *)O, ld);
  VMCX_LAUNCHED();
  return 0;
}

int slater_logderiv(int64_t B, int n_sites, int n_f, const double *ainv,
                    const uint8_t *slot, void *O, int o_dtype, int64_t ld,
                    hipStream_t st) {
  return logderiv_t<double>(B, n_sites, n_f, ainv, slot, O, o_dtype, ld, st);
}

int slater_logderiv_f32(int64_t B, int n_sites, int n_f, const float *ainv,
                        const uint8_t *slot, void *O, int o_dtype, int64_t ld,
                        hipStream_t st) {
  return logderiv_t<float>(B, n_sites, n_f, ainv, slot, O, o_dtype, ld, st);
}

// ctr[0] += inc (a draw counter advanced inside a captured graph)
__global__ void counter_add_kernel(int64_t *ctr, int64_t inc) { ctr[0] += inc; }

int counter_add(int64_t *ctr, int64_t inc, hipStream_t st) {
  VMCX_CHECK_ARG(ctr != nullptr, "null counter");
  hipLaunchKernelGGL(counter_add_kernel, dim3(1), dim3(1), 0, st, ctr, inc);
  VMCX_LAUNCHED();
  return 0;
}

// p <- p - lr * dp for FP32 parameters (the update computed in FP64)
__global__ void axpy_f32_kernel(int64_t n, double a, const double *x, float *y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)((double)y[i] + a * x[i]);
}

int sgd_update_f32(int64_t n, float *params, const double *dp, double lr, hipStream_t st) {
  VMCX_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return 0;
  VMCX_CHECK_ARG(params && dp, "null pointer");
  hipLaunchKernelGGL(axpy_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                     -lr, dp, params);
  VMCX_LAUNCHED();
  return 0;
}

}  // namespace vmcx
