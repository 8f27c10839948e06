// Chi-truncated boundary-MPS contraction of fPEPS / NN-fTNS amplitudes on
// gfx950: the reference's fPEPS_Model_GPU.amplitude with max_bond = chi
// (GPU/models/pureTNS.py:81-146: quimb contract_boundary_from_xmin over rows
// 0 .. Lx/2-1 and from_xmax over rows Lx-1 .. Lx/2, mode "mps", cutoff 0,
// canonize, then the two boundaries contracted exactly).
//
// Signed slices S_s[l,u,d,r] as in fpeps.hip (the sign convention of that
// file).  Each side is a boundary MPS B_y[a][v][b] along the row (a, b:
// boundary bonds, v: the open vertical leg).  Absorbing a row multiplies
// the bonds by the PEPS bond D:
//   A_y[(a,l)][v][(b,r)] = sum_u B_y[a][u][b] S_y[l][u][v][r]
// (u: the slice's leg facing the boundary, v: its open leg).  When a bond
// then exceeds chi, the MPS is truncated left to right onto the chi
// leading Schmidt vectors of each cut (the result of canonizing from the
// right and truncating by SVD from the left) by the density-matrix form:
//   right Grams  R_y = sum_v A_v R_{y+1} A_v^T   (y = Ly-1 .. 1, R_Ly = 1),
//   then for y = 0 .. Ly-2, with P the projection carried from the left:
//     C[(k,v)][(b,r)] = sum_{(a,l)} P[k][(a,l)] A_y[(a,l)][v][(b,r)],
//     rho = C R_{y+1} C^T, its chi leading eigenvectors U,
//     B'_y[k][v][j] = U[(k,v)][j],  P <- U^T C;
//   and B'_{Ly-1} = C.
// The tensors are Z2-graded, so rho is block diagonal in the parity of
// (k, v) (par(k) from the sector each kept eigenvector came from); each
// block is diagonalised by cyclic two-sided Jacobi in LDS (parallel
// round-robin ordering, every rotation of a round applied at once as 2 x 2
// block updates J^T rho J).  Norms are carried as FP64 log scales.
//
// One workgroup of 256 threads per (configuration, side); the Grams and the
// boundary tensors live in a per-slot workspace in HBM, the 64 x 64 working
// matrices in LDS.  A second kernel joins the two boundaries (one wave per
// configuration).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kernels.h"

namespace vmcx {

namespace {

constexpr int BT = 256;         // threads per (configuration, side)
constexpr int BN = 64;          // largest absorbed bond (chi * D)
constexpr int BLD = 65;         // LDS row stride of the BN x BN buffers
constexpr int BMAX_NS = 64;     // sites
constexpr int BMAX_LY = 16;
constexpr int BMAX_TS = 2048;   // largest stored tensor cb * D * cb
constexpr int BMAX_CHI = 32;
constexpr int BMAX_SWEEPS = 30;
// rho is scaled to a largest diagonal entry of 1; off-diagonal entries below
// this floor are not rotated away (singular values below ~1e-9 of the
// largest, which move an amplitude by about that much relative)
constexpr double BM_ROT_FLOOR = 1e-18;
// int64 header words per boundary buffer: [0] log scale (double bits),
// [1 .. Ly+1] bond dimensions, [Ly+2 .. 2Ly+2] bond parity masks
constexpr int BHDR = 40;

struct BmGeo {
  int Lx, Ly, D, N, chi, cb;    // cb = max(chi, D): bound of a stored bond
  int nbuf;                     // boundaries kept per side: 1 (amplitude), Lx - 1 (every
                                // row's environments) or max(h, Lx - h) (half, the cache)
  int half;                     // buffered sides stop at the middle (the sweep cache)
  const int32_t *dmap;
  const double *theta;
  int64_t theta_stride;
  int64_t slot;                 // doubles per (configuration, side)
  int64_t ts;                   // doubles per stored tensor, cb * D * cb
};

__device__ __forceinline__ int bqpar(int k, int d) { return d == 1 ? 0 : (k >= (d >> 1)); }
__device__ __forceinline__ int hdim(const BmGeo &g, int y) {
  return (y == 0 || y == g.Ly) ? 1 : g.D;
}

struct BmCfg {
  int p[BMAX_NS], pr[BMAX_NS], doff[BMAX_NS];
};

__device__ void bm_load_cfg(const BmGeo &g, const int64_t *c, BmCfg &cs) {
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
    int off = 0;
    for (int s = 0; s < g.N; ++s) {
      const int x = s / g.Ly, y = s - x * g.Ly;
      const int dl = y == 0 ? 1 : g.D, du = x == 0 ? 1 : g.D;
      const int dd = x == g.Lx - 1 ? 1 : g.D, dr = y == g.Ly - 1 ? 1 : g.D;
      cs.doff[s] = off;
      off += 4 * dl * du * dd * dr;
    }
  }
  __syncthreads();
}

// signed slice of site (x, y) as dst[l][f][o][r]: f the leg facing the
// boundary (up from the top side, down from the bottom), o the open leg
__device__ void bm_slice(const BmGeo &g, const BmCfg &cs, const double *__restrict__ th, int x,
                         int y, bool down, double *dst) {
  const int dl = y == 0 ? 1 : g.D, du = x == 0 ? 1 : g.D;
  const int dd = x == g.Lx - 1 ? 1 : g.D, dr = y == g.Ly - 1 ? 1 : g.D;
  const int s = x * g.Ly + y;
  const int n = dl * du * dd * dr;
  const int base = cs.doff[s] + cs.p[s] * n;
  const int pr = cs.pr[s];
  for (int idx = threadIdx.x; idx < n; idx += BT) {
    const int r = idx % dr, d = (idx / dr) % dd, u = (idx / (dr * dd)) % du;
    const int l = idx / (dr * dd * du);
    const int k = g.dmap[base + idx];
    double v = k >= 0 ? th[k] : 0.0;
    if ((bqpar(d, dd) * (bqpar(r, dr) + pr)) & 1) v = -v;
    if (down)
      dst[((l * du + u) * dd + d) * dr + r] = v;
    else
      dst[((l * dd + d) * du + u) * dr + r] = v;
  }
}

__device__ double bm_max(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = red[0];
  for (int i = 1; i < BT / 64; ++i) m = fmax(m, red[i]);
  __syncthreads();
  return m;
}

template <int NB>
struct BmShT {
  static constexpr int LD = NB + 1;   // LDS row stride of the NB x NB buffers
  double R[NB * LD];        // right Gram; then rho
  double A[NB * LD];        // A_v; then C
  double T[NB * LD];        // A_v R; then W, G, V
  // the boundary tensor of the current site (na x D x nb) and the projection
  // carried from the left (kept-bond rows): bonds never exceed NB / 2 here
  // (NB = 32 holds cb * D <= 32, so bonds <= 16 for D >= 2 and stay 1 for
  // D = 1; NB = 64 holds chi <= 32)
  double Bt[NB * (NB / 2)];
  double Sl[256];           // its slice [l][f][o][r]
  double P[(NB / 2) * NB];
  double2 jcs[NB / 2 + 1];  // rotation (c, s) of each pair slot: one 16-byte load
  double jt[NB / 2 + 1];
  int jpq[NB / 2 + 1];      // its rows p | q << 8
  int perm[NB], pos[NB], sel[NB];
  int kpar[NB];             // parity of the current left bond's indices
  int bd[BMAX_LY + 1], nbd[BMAX_LY + 1];
  uint64_t bpm[BMAX_LY + 1], npm[BMAX_LY + 1];   // bond parity masks
  double red[BT / 64];
  int sec[4];               // m0, m1, n0, n1
  int rot;
};

// B_y (dims na x D x nb, stride cb) from the slot into Bt
__device__ void bm_load_tensor(const BmGeo &g, const double *slot, int y, int na, int nb,
                               double *Bt) {
  const double *src = slot + (int64_t)y * g.ts;
  const int n = na * g.D * nb;
  for (int idx = threadIdx.x; idx < n; idx += BT) {
    const int b = idx % nb, av = idx / nb;
    Bt[idx] = src[av * g.cb + b];
  }
}

// Bt-shaped (na x D x nb, dense) LDS tensor scaled by 1 / max into the slot;
// returns ln max (-inf for a zero tensor)
__device__ double bm_store_scaled(const BmGeo &g, double *slot, int y, int na, int nb,
                                  const double *src, double *red) {
  const int n = na * g.D * nb;
  double m = 0.0;
  for (int idx = threadIdx.x; idx < n; idx += BT) m = fmax(m, fabs(src[idx]));
  m = bm_max(m, red);
  if (!(m > 0.0) || !std::isfinite(m)) return -INFINITY;
  const double inv = 1.0 / m;
  double *dst = slot + (int64_t)y * g.ts;
  for (int idx = threadIdx.x; idx < n; idx += BT) {
    const int b = idx % nb, av = idx / nb;
    dst[av * g.cb + b] = src[idx] * inv;
  }
  return log(m);
}

// out (n x m, stride LD) = X (n x K, stride LD) * Y (K x m, stride LD); TS x TS
// tiles on a 16 x 16 thread grid (TS = 2 covers 32 x 32, TS = 4 covers 64 x 64)
template <int LD, int TS>
__device__ void bm_mm(int n, int m, int K, const double *X, const double *Y, double *out) {
  const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
  const int i0 = TS * ti, j0 = TS * tj;
  if (i0 >= n || j0 >= m) return;
  double acc[TS][TS] = {};
  for (int k = 0; k < K; ++k) {
    double xa[TS], yb[TS];
#pragma unroll
    for (int q = 0; q < TS; ++q) {
      xa[q] = (i0 + q < n) ? X[(i0 + q) * LD + k] : 0.0;
      yb[q] = (j0 + q < m) ? Y[k * LD + j0 + q] : 0.0;
    }
#pragma unroll
    for (int a = 0; a < TS; ++a)
#pragma unroll
      for (int b = 0; b < TS; ++b) acc[a][b] = fma(xa[a], yb[b], acc[a][b]);
  }
#pragma unroll
  for (int a = 0; a < TS; ++a)
#pragma unroll
    for (int b = 0; b < TS; ++b)
      if (i0 + a < n && j0 + b < m) out[(i0 + a) * LD + j0 + b] = acc[a][b];
}

// cyclic Jacobi on the two diagonal blocks [0, m0) and [m0, m0 + m1) of the
// symmetric S.R, kept as its upper triangle (entry (a, b) at
// [min(a,b)][max(a,b)], so that rounding cannot make it asymmetric: a skew
// residue would survive every rotation and keep the relative stopping test
// from being met); block sizes even, padding rows / columns zero;
// eigenvectors into S.T
template <int NB>
__device__ void bm_jacobi(BmShT<NB> &S, int m0, int m1) {
  constexpr int LD = BmShT<NB>::LD;
  const int tid = threadIdx.x;
  const int np0 = m0 >> 1, np1 = m1 >> 1, npr = np0 + np1;
  const int mt = m0 + m1;
  for (int idx = tid; idx < mt * mt; idx += BT) {
    const int i = idx / mt, j = idx - i * mt;
    S.T[i * LD + j] = i == j ? 1.0 : 0.0;
  }
  const int rounds = (m0 > m1 ? m0 : m1) - 1;
  if (rounds <= 0) {
    __syncthreads();
    return;
  }
  // the work of every thread is the same in every round (only the pairing
  // changes), so its items are decoded once here, not per round:
  // phase 1: the pair this thread rotates (block b, index i, size mb)
  const int pb = tid >= np0, pi_ = pb ? tid - np0 : tid;
  const int pmb = pb ? m1 : m0, poff = pb ? m0 : 0;
  // phase 2a: pair blocks (i1 <= i2) of the upper triangle, compacted
  constexpr int NBK = ((NB / 2) * (NB / 2 + 1) / 2 + BT - 1) / BT;
  const int nt0 = np0 * (np0 + 1) / 2, nt1 = np1 * (np1 + 1) / 2;
  int bk1[NBK], bk2[NBK];
#pragma unroll
  for (int k = 0; k < NBK; ++k) {
    int wi = tid + k * BT;
    bk1[k] = -1;
    bk2[k] = -1;
    if (wi < nt0 + nt1) {
      const int b = wi >= nt0;
      const int np = b ? np1 : np0, base = b ? np0 : 0;
      if (b) wi -= nt0;
      int i1 = 0;
      while (wi >= np - i1) {   // row i1 of the triangle holds np - i1 items
        wi -= np - i1;
        ++i1;
      }
      bk1[k] = base + i1;
      bk2[k] = base + i1 + wi;
    }
  }
  // phase 2b: rows of V and the pair whose columns they rotate
  constexpr int NV = (NB * NB / 2 + BT - 1) / BT;
  const int nv0 = m0 * np0, nv1 = m1 * np1;
  int vrow[NV], vpi[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const int wi = tid + k * BT;
    vrow[k] = -1;
    vpi[k] = 0;
    if (wi < nv0) {
      vrow[k] = wi / np0;
      vpi[k] = wi - vrow[k] * np0;
    } else if (wi < nv0 + nv1) {
      const int w2 = wi - nv0;
      vrow[k] = m0 + w2 / np1;
      vpi[k] = np0 + w2 % np1;
    }
  }
  for (int sweep = 0; sweep < BMAX_SWEEPS; ++sweep) {
    if (tid == 0) S.rot = 0;
    __syncthreads();
    int rr = 0;   // the round within a block of size pmb: r mod (pmb - 1)
    for (int r = 0; r < rounds; ++r) {
      if (tid < npr) {
        int p = 0, q = 0;
        double c = 1.0, s = 0.0, t = 0.0;
        if (pmb >= 2) {
          if (pi_ == 0) {
            p = pmb - 1;
            q = rr;
          } else {
            p = rr + pi_;
            if (p >= pmb - 1) p -= pmb - 1;
            q = rr - pi_;
            if (q < 0) q += pmb - 1;
          }
          p += poff;
          q += poff;
          const double app = S.R[p * LD + p], aqq = S.R[q * LD + q];
          const double apq = S.R[(p < q ? p : q) * LD + (p < q ? q : p)];
          // |a_pq| > max(1e-15 sqrt|a_pp a_qq|, floor), compared in squares (no square root)
          if (apq * apq > fmax(1e-30 * fabs(app * aqq), BM_ROT_FLOOR * BM_ROT_FLOOR)) {
            // t = sgn(theta) / (|theta| + sqrt(theta^2 + 1)), theta = tau / zeta, written
            // with one square root and one division (|t| <= 1); c = 1 / sqrt(1 + t^2)
            // by the hardware estimate and three Newton steps (the estimate's 2^-22
            // relative error squared three times is below the FP64 rounding)
            const double tau = aqq - app, zeta = 2.0 * apq;
            t = copysign(1.0, tau) * zeta / (fabs(tau) + sqrt(tau * tau + zeta * zeta));
            const double x = 1.0 + t * t;
            double y = __builtin_amdgcn_rsq(x);
#pragma unroll
            for (int it = 0; it < 3; ++it) y = y * fma(-0.5 * x * y, y, 1.5);
            c = y;
            s = t * c;
            S.rot = 1;
          }
          if (++rr == pmb - 1) rr = 0;
        }
        S.jpq[tid] = p | (q << 8);
        S.jcs[tid] = make_double2(c, s);
        S.jt[tid] = t;
      }
      __syncthreads();
      // rho <- J^T rho J on the upper triangle of pair blocks
#pragma unroll
      for (int k = 0; k < NBK; ++k) {
        const int i1 = bk1[k], i2 = bk2[k];
        if (i1 < 0) continue;
        const double2 cs1 = S.jcs[i1], cs2 = S.jcs[i2];
        const double s1 = cs1.y, s2 = cs2.y;
        if (s1 == 0.0 && s2 == 0.0) continue;
        const int pq1 = S.jpq[i1];
        const int p1 = pq1 & 255, q1 = pq1 >> 8;
        if (i1 == i2) {
          // the rotated pair itself: its off-diagonal entry becomes zero
          const double t = S.jt[i1];
          const int lo = p1 < q1 ? p1 : q1, hi = p1 < q1 ? q1 : p1;
          const double apq = S.R[lo * LD + hi];
          S.R[p1 * LD + p1] -= t * apq;
          S.R[q1 * LD + q1] += t * apq;
          S.R[lo * LD + hi] = 0.0;
          continue;
        }
        const int pq2 = S.jpq[i2];
        const int p2 = pq2 & 255, q2 = pq2 >> 8;
        const double c1 = cs1.x, c2 = cs2.x;
        // entries of distinct pairs: never on the diagonal
        const int e0 = p1 < p2 ? p1 * LD + p2 : p2 * LD + p1;
        const int e1 = p1 < q2 ? p1 * LD + q2 : q2 * LD + p1;
        const int e2 = q1 < p2 ? q1 * LD + p2 : p2 * LD + q1;
        const int e3 = q1 < q2 ? q1 * LD + q2 : q2 * LD + q1;
        const double xpp = S.R[e0], xpq = S.R[e1], xqp = S.R[e2], xqq = S.R[e3];
        const double ypp = c1 * xpp - s1 * xqp, ypq = c1 * xpq - s1 * xqq;
        const double yqp = s1 * xpp + c1 * xqp, yqq = s1 * xpq + c1 * xqq;
        S.R[e0] = ypp * c2 - ypq * s2;
        S.R[e1] = ypp * s2 + ypq * c2;
        S.R[e2] = yqp * c2 - yqq * s2;
        S.R[e3] = yqp * s2 + yqq * c2;
      }
      // V <- V J over the rows of each block (loads issued before the stores)
      {
        int ip[NV], iq[NV];
        double vp[NV], vq[NV], cc[NV], ss[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          const bool on = vrow[k] >= 0;
          const double2 cs = on ? S.jcs[vpi[k]] : make_double2(1.0, 0.0);
          const int pq = on ? S.jpq[vpi[k]] : 0;
          ss[k] = cs.y;
          cc[k] = cs.x;
          ip[k] = on ? vrow[k] * LD + (pq & 255) : 0;
          iq[k] = on ? vrow[k] * LD + (pq >> 8) : 0;
        }
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          vp[k] = S.T[ip[k]];
          vq[k] = S.T[iq[k]];
        }
#pragma unroll
        for (int k = 0; k < NV; ++k)
          if (ss[k] != 0.0) {
            S.T[ip[k]] = cc[k] * vp[k] - ss[k] * vq[k];
            S.T[iq[k]] = ss[k] * vp[k] + cc[k] * vq[k];
          }
      }
      __syncthreads();
    }
    const int any = S.rot;
    __syncthreads();
    if (!any) break;
  }
}

}  // namespace

// the side steps a move between sites sa and sb (spin offset N allowed)
// changes: st >= the side's row count means unchanged
__device__ __forceinline__ void bm_move_starts(const BmGeo &g, uint32_t mv, int &st_top,
                                               int &st_bot) {
  int sa = mv & 0xFF, sb = (mv >> 8) & 0xFF;
  if (sa >= g.N) sa -= g.N;
  if (sb >= g.N) sb -= g.N;
  const int ra = sa / g.Ly, rb = sb / g.Ly, h = g.Lx / 2;
  const int rmin = ra < rb ? ra : rb, rmax = ra < rb ? rb : ra;
  st_top = rmin <= h - 1 ? rmin : h;
  st_bot = rmax >= h ? g.Lx - 1 - rmax : g.Lx - h;
}

// ------------------------------------------------------------ one side
// grid (rows, 2): blockIdx.y = 0 the top side (rows 0 .. Lx/2-1, open down
// legs), 1 the bottom side (rows Lx-1 .. Lx/2, open up legs).  With
// g.nbuf = Lx - 1 every side runs over Lx - 1 rows and keeps the boundary
// after each (buffer i: rows 0 .. i from the top, Lx-1 .. Lx-1-i from the
// bottom): the environments of every row (quimb's compute_x_environments);
// with g.nbuf = 1 it stops at the middle and works in place.
template <int NB>
__global__ void __launch_bounds__(BT, NB == 32 ? 4 : 1)
    peps_bmps_side_kernel(BmGeo g, int64_t rows, const int64_t *__restrict__ cfg,
                          const int32_t *__restrict__ flag, double *ws, int64_t *hdr,
                          int st_top, int st_bot, const double *__restrict__ cache,
                          const int64_t *__restrict__ chdr, const int32_t *__restrict__ parent,
                          const uint32_t *__restrict__ cmove) {
  constexpr int LD = BmShT<NB>::LD;
  constexpr int TS = NB / 16;   // register tile of the dense products (16 x 16 threads)
  const int64_t w = blockIdx.x;
  if (w >= rows) return;
  if (flag && !(flag[w] & 1)) return;
  const int side = blockIdx.y;
  const bool down = side == 0;
  const int tid = threadIdx.x;
  // rows 0 .. start-1 of this side are those of the cached walker (parent[w],
  // or w): resume from its boundary after them (start >= nrows: the side is
  // unchanged); per configuration from its move when cmove is given
  if (cache && cmove) bm_move_starts(g, cmove[w], st_top, st_bot);
  const int start = cache ? (down ? st_top : st_bot) : 0;
  const int64_t pw = parent ? (int64_t)parent[w] : w;
  __shared__ BmCfg cs;
  __shared__ BmShT<NB> S;
  bm_load_cfg(g, cfg + w * g.N, cs);
  const double *th = g.theta + w * g.theta_stride;
  double *slot = ws + (w * 2 + side) * g.slot;
  const int64_t bstride = (int64_t)g.Ly * g.ts;
  double *Rg = slot + g.nbuf * bstride;
  int64_t *hd = hdr + (w * 2 + side) * g.nbuf * BHDR;
  const int Ly = g.Ly, D = g.D, h = g.Lx / 2;
  const int nrows = (g.nbuf > 1 && !g.half) ? g.Lx - 1 : (down ? h : g.Lx - h);
  if (start >= nrows) return;
  const int bufd = g.nbuf > 1 ? 1 : 0;   // buffer index step per row
  const double *cslot = cache ? cache + (pw * 2 + side) * g.slot : nullptr;
  const int64_t *chd = chdr ? chdr + (pw * 2 + side) * g.nbuf * BHDR : nullptr;
  const uint64_t dmask = D == 1 ? 0ull : (((1ull << D) - 1) & ~((1ull << (D >> 1)) - 1));
  double lg = 0.0;
  bool dead = false;
  if (start > 0) {
    const int64_t *c = chd + (start - 1) * bufd * BHDR;
    if (tid <= Ly) {
      S.bd[tid] = (int)c[1 + tid];
      S.bpm[tid] = (uint64_t)c[Ly + 2 + tid];
    }
    lg = __longlong_as_double(c[0]);
    dead = std::isinf(lg);
  } else if (tid <= Ly) {
    S.bd[tid] = hdim(g, tid);
    S.bpm[tid] = hdim(g, tid) == 1 ? 0ull : dmask;
  }
  __syncthreads();
  // first row: B_y[l][v][r] = S[l][0][v][r]
  if (start == 0) {
    const int x = down ? 0 : g.Lx - 1;
    for (int y = 0; y < Ly && !dead; ++y) {
      bm_slice(g, cs, th, x, y, down, S.Sl);
      __syncthreads();
      const int dl = hdim(g, y), dr = hdim(g, y + 1);
      const double l2 = bm_store_scaled(g, slot, y, dl, dr, S.Sl, S.red);
      __syncthreads();
      if (std::isinf(l2)) dead = true;
      lg += l2;
    }
    if (tid <= Ly) {
      hd[1 + tid] = S.bd[tid];
      hd[Ly + 2 + tid] = (int64_t)S.bpm[tid];
    }
    if (tid == 0) hd[0] = __double_as_longlong(dead ? -INFINITY : lg);
  }
  for (int ri = start > 1 ? start : 1; ri < nrows && !dead; ++ri) {
    const int x = down ? ri : g.Lx - 1 - ri;
    // boundary after row ri - 1 (the cache's for the first resumed row) ->
    // after row ri (the same buffer when nbuf = 1)
    const double *src = (start > 0 && ri == start) ? cslot + (ri - 1) * bstride
                                                   : slot + (ri - 1) * bufd * bstride;
    double *dst = slot + ri * bufd * bstride;
    if (tid <= Ly) {
      const int hd = hdim(g, tid), b = S.bd[tid];
      S.nbd[tid] = b * hd;
      uint64_t m = 0;
      for (int a = 0; a < b; ++a)
        for (int l = 0; l < hd; ++l)
          if ((((S.bpm[tid] >> a) & 1) ^ bqpar(l, hd))) m |= 1ull << (a * hd + l);
      S.npm[tid] = m;
    }
    __syncthreads();
    int mxb = 0;
    for (int y = 0; y <= Ly; ++y) mxb = S.nbd[y] > mxb ? S.nbd[y] : mxb;
    if (mxb <= g.chi) {
      // absorb without truncation
      for (int y = 0; y < Ly && !dead; ++y) {
        const int na = S.bd[y], nb = S.bd[y + 1], dl = hdim(g, y), dr = hdim(g, y + 1);
        bm_load_tensor(g, src, y, na, nb, S.Bt);
        bm_slice(g, cs, th, x, y, down, S.Sl);
        __syncthreads();
        const int ni = na * dl, nj = nb * dr, n = ni * D * nj;
        for (int idx = tid; idx < n; idx += BT) {
          const int j = idx % nj, v = (idx / nj) % D, i = idx / (nj * D);
          const int a = i / dl, l = i - a * dl, b = j / dr, r = j - b * dr;
          double s = 0.0;
          for (int u = 0; u < D; ++u)
            s = fma(S.Bt[(a * D + u) * nb + b], S.Sl[((l * D + u) * D + v) * dr + r], s);
          S.A[idx] = s;
        }
        __syncthreads();
        const double l2 = bm_store_scaled(g, dst, y, ni, nj, S.A, S.red);
        __syncthreads();
        if (std::isinf(l2)) dead = true;
        lg += l2;
      }
      if (tid <= Ly) {
        S.bd[tid] = S.nbd[tid];
        S.bpm[tid] = S.npm[tid];
      }
      __syncthreads();
      if (tid <= Ly) {
        hd[ri * bufd * BHDR + 1 + tid] = S.bd[tid];
        hd[ri * bufd * BHDR + Ly + 2 + tid] = (int64_t)S.bpm[tid];
      }
      if (tid == 0) hd[ri * bufd * BHDR] = __double_as_longlong(dead ? -INFINITY : lg);
      continue;
    }
    // ---- backward: right Grams R_y, y = Ly-1 .. 1 (R_Ly = 1)
    if (tid == 0) S.R[0] = 1.0;
    for (int y = Ly - 1; y >= 1; --y) {
      const int na = S.bd[y], nb = S.bd[y + 1], dl = hdim(g, y), dr = hdim(g, y + 1);
      const int n = na * dl, m = nb * dr;
      bm_load_tensor(g, src, y, na, nb, S.Bt);
      bm_slice(g, cs, th, x, y, down, S.Sl);
      __syncthreads();
      const int ti = tid >> 4, tj = tid & 15, i0 = TS * ti, j0 = TS * tj;
      double acc[TS][TS] = {};
      for (int v = 0; v < D; ++v) {
        for (int idx = tid; idx < n * m; idx += BT) {
          const int i = idx / m, j = idx - i * m;
          const int a = i / dl, l = i - a * dl, b = j / dr, r = j - b * dr;
          double s = 0.0;
          for (int u = 0; u < D; ++u)
            s = fma(S.Bt[(a * D + u) * nb + b], S.Sl[((l * D + u) * D + v) * dr + r], s);
          S.A[i * LD + j] = s;
        }
        __syncthreads();
        bm_mm<LD, TS>(n, m, m, S.A, S.R, S.T);
        __syncthreads();
        if (i0 < n && j0 < n) {
          for (int k = 0; k < m; ++k) {
            double xa[TS], yb[TS];
#pragma unroll
            for (int q = 0; q < TS; ++q) {
              xa[q] = (i0 + q < n) ? S.T[(i0 + q) * LD + k] : 0.0;
              yb[q] = (j0 + q < n) ? S.A[(j0 + q) * LD + k] : 0.0;
            }
#pragma unroll
            for (int a = 0; a < TS; ++a)
#pragma unroll
              for (int b = 0; b < TS; ++b) acc[a][b] = fma(xa[a], yb[b], acc[a][b]);
          }
        }
        __syncthreads();
      }
      double mx = 0.0;
#pragma unroll
      for (int a = 0; a < TS; ++a)
#pragma unroll
        for (int b = 0; b < TS; ++b) mx = fmax(mx, fabs(acc[a][b]));
      mx = bm_max(mx, S.red);
      const double inv = mx > 0.0 ? 1.0 / mx : 0.0;
      double *Rd = Rg + (int64_t)y * BN * BN;
#pragma unroll
      for (int a = 0; a < TS; ++a)
#pragma unroll
        for (int b = 0; b < TS; ++b)
          if (i0 + a < n && j0 + b < n) {
            const double v = acc[a][b] * inv;
            S.R[(i0 + a) * LD + j0 + b] = v;
            Rd[(i0 + a) * BN + j0 + b] = v;
          }
      __syncthreads();
    }
    // ---- forward: truncation left to right
    if (tid == 0) {
      S.P[0] = 1.0;
      S.kpar[0] = 0;
      S.nbd[0] = 1;
      S.npm[0] = 0;
    }
    __syncthreads();
    int k = 1;
    for (int y = 0; y < Ly && !dead; ++y) {
      const int na = S.bd[y], nb = S.bd[y + 1], dl = hdim(g, y), dr = hdim(g, y + 1);
      const int m = nb * dr, nC = k * D;
      bm_load_tensor(g, src, y, na, nb, S.Bt);
      bm_slice(g, cs, th, x, y, down, S.Sl);
      __syncthreads();
      // W[kk][l][u][b] = sum_a P[kk][(a,l)] B[a][u][b]  (into T)
      const int nw = k * dl * D * nb;
      for (int idx = tid; idx < nw; idx += BT) {
        const int b = idx % nb, u = (idx / nb) % D, l = (idx / (nb * D)) % dl;
        const int kk = idx / (nb * D * dl);
        double s = 0.0;
        for (int a = 0; a < na; ++a)
          s = fma(S.P[kk * NB + a * dl + l], S.Bt[(a * D + u) * nb + b], s);
        S.T[idx] = s;
      }
      __syncthreads();
      // C[(kk,v)][(b,r)] = sum_{l,u} W[kk][l][u][b] S[l][u][v][r]  (into A)
      for (int idx = tid; idx < nC * m; idx += BT) {
        const int j = idx % m, i = idx / m;
        const int kk = i / D, v = i - kk * D, b = j / dr, r = j - b * dr;
        double s = 0.0;
        for (int l = 0; l < dl; ++l)
          for (int u = 0; u < D; ++u)
            s = fma(S.T[((kk * dl + l) * D + u) * nb + b], S.Sl[((l * D + u) * D + v) * dr + r],
                    s);
        S.A[i * LD + j] = s;
      }
      __syncthreads();
      if (y == Ly - 1) {
        // last site: B'[kk][v][0] = C[(kk,v)][0]
        for (int idx = tid; idx < nC; idx += BT) S.T[idx] = S.A[idx * LD];
        __syncthreads();
        const double l2 = bm_store_scaled(g, dst, y, k, 1, S.T, S.red);
        __syncthreads();
        if (std::isinf(l2)) dead = true;
        lg += l2;
        if (tid == 0) {
          S.nbd[Ly] = 1;
          S.npm[Ly] = 0;
        }
        break;
      }
      double *dsty = dst + (int64_t)y * g.ts;
      int kn;
      if (nC <= g.chi) {
        // no truncation at this bond: the identity basis, P = C
        kn = nC;
        for (int idx = tid; idx < k * D * kn; idx += BT) {
          const int j = idx % kn, kv = idx / kn;
          dsty[kv * g.cb + j] = j == kv ? 1.0 : 0.0;
        }
        for (int idx = tid; idx < kn * m; idx += BT) {
          const int j = idx % m, i = idx / m;
          S.P[i * NB + j] = S.A[i * LD + j];
        }
        __syncthreads();
        if (tid < kn) S.sel[tid] = S.kpar[tid / D] ^ bqpar(tid % D, D);
        __syncthreads();
        if (tid < kn) S.kpar[tid] = S.sel[tid];
      } else {
        kn = g.chi;
        // R_{y+1} from the slot
        const double *Rs = Rg + (int64_t)(y + 1) * BN * BN;
        for (int idx = tid; idx < m * m; idx += BT) {
          const int i = idx / m, j = idx - i * m;
          S.R[i * LD + j] = Rs[i * BN + j];
        }
        if (tid == 0) {
          int n0 = 0;
          for (int i = 0; i < nC; ++i) n0 += !(S.kpar[i / D] ^ bqpar(i % D, D));
          const int n1 = nC - n0, m0 = n0 + (n0 & 1), m1 = n1 + (n1 & 1);
          int c0 = 0, c1 = m0;
          for (int i = 0; i < nC; ++i) {
            const int sct = S.kpar[i / D] ^ bqpar(i % D, D);
            const int pi = sct ? c1++ : c0++;
            S.perm[pi] = i;
            S.pos[i] = pi;
          }
          if (n0 & 1) S.perm[m0 - 1] = -1;
          if (n1 & 1) S.perm[m0 + m1 - 1] = -1;
          S.sec[0] = m0;
          S.sec[1] = m1;
        }
        __syncthreads();
        const int m0 = S.sec[0], m1 = S.sec[1], mt = m0 + m1;
        // G = C R (into T)
        bm_mm<LD, TS>(nC, m, m, S.A, S.R, S.T);
        __syncthreads();
        // rho (permuted, block diagonal) = G C^T (into R)
        for (int idx = tid; idx < mt * mt; idx += BT) {
          const int p1 = idx / mt, p2 = idx - p1 * mt;
          const int i1 = S.perm[p1], i2 = S.perm[p2];
          double s = 0.0;
          if (p1 <= p2 && i1 >= 0 && i2 >= 0 && ((p1 < m0) == (p2 < m0)))
            for (int j = 0; j < m; ++j) s = fma(S.T[i1 * LD + j], S.A[i2 * LD + j], s);
          S.R[p1 * LD + p2] = s;     // upper triangle only (the rest unused)
        }
        __syncthreads();
        double dmx = 0.0;
        for (int i = tid; i < mt; i += BT) dmx = fmax(dmx, fabs(S.R[i * LD + i]));
        dmx = bm_max(dmx, S.red);
        if (!(dmx > 0.0)) {
          dead = true;
          break;
        }
        const double dinv = 1.0 / dmx;
        for (int idx = tid; idx < mt * mt; idx += BT) {
          const int p1 = idx / mt, p2 = idx - p1 * mt;
          S.R[p1 * LD + p2] *= dinv;
        }
        __syncthreads();
        bm_jacobi<NB>(S, m0, m1);
        // the chi leading eigenvalues (ties by position), padding excluded
        if (tid < mt && S.perm[tid] >= 0) {
          const double lam = S.R[tid * LD + tid];
          int rank = 0;
          for (int q = 0; q < mt; ++q) {
            if (S.perm[q] < 0) continue;
            const double lq = S.R[q * LD + q];
            rank += (lq > lam) || (lq == lam && q < tid);
          }
          if (rank < kn) S.sel[rank] = tid;
        }
        __syncthreads();
        // B'[kk][v][j] = V[pos(kk,v)][sel j]
        for (int idx = tid; idx < nC * kn; idx += BT) {
          const int j = idx % kn, kv = idx / kn;
          dsty[kv * g.cb + j] = S.T[S.pos[kv] * LD + S.sel[j]];
        }
        // P'[j][(b,r)] = sum_pi V[pi][sel j] C[perm pi][(b,r)]
        for (int idx = tid; idx < kn * m; idx += BT) {
          const int jj = idx % m, j = idx / m;
          const int sj = S.sel[j];
          const int lo = sj < m0 ? 0 : m0, hi = sj < m0 ? m0 : mt;
          double s = 0.0;
          for (int pi = lo; pi < hi; ++pi) {
            const int i = S.perm[pi];
            if (i >= 0) s = fma(S.T[pi * LD + sj], S.A[i * LD + jj], s);
          }
          S.P[j * NB + jj] = s;
        }
        __syncthreads();
        if (tid < kn) S.kpar[tid] = S.sel[tid] >= m0;
      }
      __syncthreads();
      // renormalise P
      double pm = 0.0;
      for (int idx = tid; idx < kn * m; idx += BT) {
        const int j = idx % m, i = idx / m;
        pm = fmax(pm, fabs(S.P[i * NB + j]));
      }
      pm = bm_max(pm, S.red);
      if (!(pm > 0.0) || !std::isfinite(pm)) {
        dead = true;
        break;
      }
      const double pinv = 1.0 / pm;
      for (int idx = tid; idx < kn * m; idx += BT) {
        const int j = idx % m, i = idx / m;
        S.P[i * NB + j] *= pinv;
      }
      lg += log(pm);
      if (tid == 0) {
        S.nbd[y + 1] = kn;
        uint64_t msk = 0;
        for (int j = 0; j < kn; ++j) msk |= (uint64_t)S.kpar[j] << j;
        S.npm[y + 1] = msk;
      }
      __syncthreads();
      k = kn;
    }
    __syncthreads();
    if (tid <= Ly) {
      S.bd[tid] = S.nbd[tid];
      S.bpm[tid] = S.npm[tid];
    }
    __syncthreads();
    if (tid <= Ly) {
      hd[ri * bufd * BHDR + 1 + tid] = S.bd[tid];
      hd[ri * bufd * BHDR + Ly + 2 + tid] = (int64_t)S.bpm[tid];
    }
    if (tid == 0) hd[ri * bufd * BHDR] = __double_as_longlong(dead ? -INFINITY : lg);
  }
  // buffers never reached (a vanishing boundary): mark them zero
  if (dead && tid == 0)
    for (int i = start; i < nrows; ++i) hd[i * bufd * BHDR] = __double_as_longlong(-INFINITY);
}

// ------------------------------------------------------------ join
// the amplitude from the top boundary in buffer bt and the bottom one in bb
__global__ void __launch_bounds__(64)
    peps_bmps_join_kernel(BmGeo g, int64_t rows, const int64_t *__restrict__ cfg,
                          const int32_t *__restrict__ flag, const double *__restrict__ ws,
                          const int64_t *__restrict__ hdr, int bt, int bb, double *sign,
                          double *logabs, int ct, int cb_, const double *__restrict__ cache,
                          const int64_t *__restrict__ chdr, const int32_t *__restrict__ parent,
                          const uint32_t *__restrict__ cmove) {
  const int64_t w = blockIdx.x;
  if (w >= rows) return;
  if (flag && !(flag[w] & 1)) return;
  const int tid = threadIdx.x;
  __shared__ double E[BMAX_CHI * BMAX_CHI], X[BMAX_TS];
  // a side taken from the cache (ct / cb_ set: unchanged since the walker
  // parent[w], or w; per configuration from its move when cmove is given)
  if (cache && cmove) {
    int s0, s1;
    bm_move_starts(g, cmove[w], s0, s1);
    ct = s0 >= g.Lx / 2;
    cb_ = s1 >= g.Lx - g.Lx / 2;
  }
  const int64_t pw = parent ? (int64_t)parent[w] : w;
  const int64_t *ht = ct ? chdr + ((pw * 2) * g.nbuf + bt) * BHDR
                         : hdr + ((w * 2) * g.nbuf + bt) * BHDR;
  const int64_t *hb = cb_ ? chdr + ((pw * 2 + 1) * g.nbuf + bb) * BHDR
                          : hdr + ((w * 2 + 1) * g.nbuf + bb) * BHDR;
  const double lt = __longlong_as_double(ht[0]), lb = __longlong_as_double(hb[0]);
  const double lo = log(1e-45);
  int P = 0;
  for (int s = 0; s < g.N; ++s) {
    const int v = (int)cfg[w * g.N + s];
    P += (v == 1 || v == 2);
  }
  if (std::isinf(lt) || std::isinf(lb)) {
    if (tid == 0) {
      sign[w] = 0.0;
      logabs[w] = lo;
    }
    return;
  }
  const int64_t bstride = (int64_t)g.Ly * g.ts;
  const double *Tt = ct ? cache + (pw * 2) * g.slot + bt * bstride
                        : ws + (w * 2) * g.slot + bt * bstride;
  const double *Tb = cb_ ? cache + (pw * 2 + 1) * g.slot + bb * bstride
                         : ws + (w * 2 + 1) * g.slot + bb * bstride;
  const int D = g.D, cb = g.cb;
  if (tid == 0) E[0] = 1.0;
  __syncthreads();
  double lsc = 0.0;
  for (int y = 0; y < g.Ly; ++y) {
    const int ta = (int)ht[1 + y], tb = (int)ht[2 + y], ba = (int)hb[1 + y], bb = (int)hb[2 + y];
    const double *T = Tt + (int64_t)y * g.ts, *Bo = Tb + (int64_t)y * g.ts;
    // X[a'][v][b] = sum_a E[a][a'] T[a][v][b]
    for (int idx = tid; idx < ba * D * tb; idx += 64) {
      const int b = idx % tb, v = (idx / tb) % D, ap = idx / (tb * D);
      double s = 0.0;
      for (int a = 0; a < ta; ++a) s = fma(E[a * BMAX_CHI + ap], T[(a * D + v) * cb + b], s);
      X[idx] = s;
    }
    __syncthreads();
    // E'[b][b'] = sum_{a',v} X[a'][v][b] Bo[a'][v][b']
    double mx = 0.0;
    for (int idx = tid; idx < tb * bb; idx += 64) {
      const int bp = idx % bb, b = idx / bb;
      double s = 0.0;
      for (int ap = 0; ap < ba; ++ap)
        for (int v = 0; v < D; ++v) s = fma(X[(ap * D + v) * tb + b], Bo[(ap * D + v) * cb + bp], s);
      E[b * BMAX_CHI + bp] = s;
      mx = fmax(mx, fabs(s));
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    __syncthreads();
    if (!(mx > 0.0) || !std::isfinite(mx)) {
      if (tid == 0) {
        sign[w] = 0.0;
        logabs[w] = lo;
      }
      return;
    }
    const double inv = 1.0 / mx;
    for (int idx = tid; idx < tb * bb; idx += 64) {
      const int bp = idx % bb, b = idx / bb;
      E[b * BMAX_CHI + bp] *= inv;
    }
    lsc += log(mx);
    __syncthreads();
  }
  if (tid == 0) {
    const double v = E[0] * (((P * (P - 1) / 2) & 1) ? -1.0 : 1.0);
    if (v == 0.0 || !std::isfinite(v)) {
      sign[w] = 0.0;
      logabs[w] = lo;
    } else {
      sign[w] = v > 0.0 ? 1.0 : -1.0;
      logabs[w] = fmax(log(fabs(v)) + lsc + lt + lb, lo);
    }
  }
}

// ------------------------------------------------------------ cheap gradient
// d ln|psi| / d theta of the sites of row r with the boundary environments of
// the row held fixed (the reference's cheap gradient, pureTNS.py:2230-2334):
// Z_r = <top_r| row r |bot_r>, top_r the boundary of rows < r (buffer r-1 of
// the top side), bot_r of rows > r (buffer Lx-2-r of the bottom side); the
// right strip environments R_y[b][r][b'] first (kept in renv), then a left
// sweep with L_y[a][l][a'] that forms
//   dZ/dS_y[l][u][d][r] = sum L_y[a][l][a'] T_y[a][u][b] Bo_y[a'][d][b'] R_{y+1}[b][r][b'],
// divided by psi (sign, logabs of the join) and scattered to theta's entries
// of the configuration's local states (G rows zeroed beforehand).
struct GrSh {
  double T[BMAX_TS], Bo[BMAX_TS], L[BMAX_TS], R[BMAX_TS];
  double X[BN * BN], Y[BN * BN];
  double Sl[256], dS[256];
  double red[BT / 64];
  int bt[BMAX_LY + 1], bb[BMAX_LY + 1];
};

__global__ void __launch_bounds__(BT)
    peps_bmps_grad_kernel(BmGeo g, int64_t rows, const int64_t *__restrict__ cfg,
                          const int32_t *__restrict__ flag, const double *__restrict__ ws,
                          const int64_t *__restrict__ hdr, const double *__restrict__ sign,
                          const double *__restrict__ logabs, double *renv, double *G,
                          int64_t ldg) {
  const int64_t w = blockIdx.x;
  const int r = blockIdx.y;
  if (w >= rows) return;
  if (flag && !(flag[w] & 1)) return;
  if (sign[w] == 0.0) return;
  const int tid = threadIdx.x;
  __shared__ BmCfg cs;
  __shared__ GrSh S;
  bm_load_cfg(g, cfg + w * g.N, cs);
  const double *th = g.theta + w * g.theta_stride;
  const int Lx = g.Lx, Ly = g.Ly, D = g.D;
  const int64_t bstride = (int64_t)Ly * g.ts;
  const bool ttriv = r == 0, btriv = r == Lx - 1;
  const double *tsrc = ws + (w * 2) * g.slot + (ttriv ? 0 : (r - 1) * bstride);
  const double *bsrc = ws + (w * 2 + 1) * g.slot + (btriv ? 0 : (Lx - 2 - r) * bstride);
  const int64_t *th_ = hdr + ((w * 2) * g.nbuf + (ttriv ? 0 : r - 1)) * BHDR;
  const int64_t *hb_ = hdr + ((w * 2 + 1) * g.nbuf + (btriv ? 0 : Lx - 2 - r)) * BHDR;
  const double lt = ttriv ? 0.0 : __longlong_as_double(th_[0]);
  const double lb = btriv ? 0.0 : __longlong_as_double(hb_[0]);
  if (std::isinf(lt) || std::isinf(lb)) return;
  if (tid <= Ly) {
    S.bt[tid] = ttriv ? 1 : (int)th_[1 + tid];
    S.bb[tid] = btriv ? 1 : (int)hb_[1 + tid];
  }
  int P = 0;
  for (int q = 0; q < g.N; ++q) P += (cs.p[q] == 1 || cs.p[q] == 2);
  const double gsign = ((P * (P - 1) / 2) & 1) ? -1.0 : 1.0;
  const int du = r == 0 ? 1 : D, dd = r == Lx - 1 ? 1 : D;
  double *rs = renv + (w * Lx + r) * ((int64_t)(Ly + 1) * g.ts + Ly + 1);
  double *rlog = rs + (int64_t)(Ly + 1) * g.ts;
  __syncthreads();
  // load T_y (ta x du x tb) and Bo_y (ba x dd x bb) densely
  auto load_env = [&](int y) {
    const int ta = S.bt[y], tb = S.bt[y + 1], ba = S.bb[y], bb = S.bb[y + 1];
    for (int idx = tid; idx < ta * du * tb; idx += BT) {
      const int b = idx % tb, au = idx / tb;
      S.T[idx] = ttriv ? 1.0 : tsrc[(int64_t)y * g.ts + au * g.cb + b];
    }
    for (int idx = tid; idx < ba * dd * bb; idx += BT) {
      const int b = idx % bb, ad = idx / bb;
      S.Bo[idx] = btriv ? 1.0 : bsrc[(int64_t)y * g.ts + ad * g.cb + b];
    }
  };
  // ---- right strip environments R_y, y = Ly-1 .. 1 (R_Ly = 1)
  if (tid == 0) {
    S.R[0] = 1.0;
    rlog[Ly] = 0.0;
  }
  double lr = 0.0;
  for (int y = Ly - 1; y >= 1; --y) {
    const int ta = S.bt[y], tb = S.bt[y + 1], ba = S.bb[y], bb = S.bb[y + 1];
    const int dl = hdim(g, y), dr = hdim(g, y + 1);
    load_env(y);
    bm_slice(g, cs, th, r, y, true, S.Sl);
    __syncthreads();
    // Z1[a][u][rr][b'] = sum_b T[a][u][b] R[b][rr][b']
    for (int idx = tid; idx < ta * du * dr * bb; idx += BT) {
      const int bp = idx % bb, rr = (idx / bb) % dr, au = idx / (bb * dr);
      double v = 0.0;
      for (int b = 0; b < tb; ++b) v = fma(S.T[au * tb + b], S.R[(b * dr + rr) * bb + bp], v);
      S.X[idx] = v;
    }
    __syncthreads();
    // Z2[a][l][d][b'] = sum_{u,rr} Z1[a][u][rr][b'] S[l][u][d][rr]
    for (int idx = tid; idx < ta * dl * dd * bb; idx += BT) {
      const int bp = idx % bb, d = (idx / bb) % dd, l = (idx / (bb * dd)) % dl;
      const int a = idx / (bb * dd * dl);
      double v = 0.0;
      for (int u = 0; u < du; ++u)
        for (int rr = 0; rr < dr; ++rr)
          v = fma(S.X[((a * du + u) * dr + rr) * bb + bp], S.Sl[((l * du + u) * dd + d) * dr + rr],
                  v);
      S.Y[idx] = v;
    }
    __syncthreads();
    // R_y[a][l][a'] = sum_{d,b'} Z2[a][l][d][b'] Bo[a'][d][b']
    double mx = 0.0;
    const int nr = ta * dl * ba;
    for (int idx = tid; idx < nr; idx += BT) {
      const int ap = idx % ba, al = idx / ba;
      double v = 0.0;
      for (int d = 0; d < dd; ++d)
        for (int bp = 0; bp < bb; ++bp)
          v = fma(S.Y[(al * dd + d) * bb + bp], S.Bo[(ap * dd + d) * bb + bp], v);
      S.L[idx] = v;
      mx = fmax(mx, fabs(v));
    }
    mx = bm_max(mx, S.red);
    if (!(mx > 0.0) || !std::isfinite(mx)) return;   // uniform: a vanishing strip
    const double inv = 1.0 / mx;
    lr += log(mx);
    for (int idx = tid; idx < nr; idx += BT) {
      const double v = S.L[idx] * inv;
      S.R[idx] = v;
      rs[(int64_t)y * g.ts + idx] = v;
    }
    if (tid == 0) rlog[y] = lr;
    __syncthreads();
  }
  // ---- left sweep with the derivative of every site
  if (tid == 0) S.L[0] = 1.0;
  double ll = 0.0;
  const double base = lt + lb - logabs[w];
  const double sg = sign[w] * gsign;
  for (int y = 0; y < Ly; ++y) {
    const int ta = S.bt[y], tb = S.bt[y + 1], ba = S.bb[y], bb = S.bb[y + 1];
    const int dl = hdim(g, y), dr = hdim(g, y + 1);
    load_env(y);
    bm_slice(g, cs, th, r, y, true, S.Sl);
    const double lry = y == Ly - 1 ? 0.0 : rlog[y + 1];
    if (y == Ly - 1) {
      if (tid == 0) S.R[0] = 1.0;
    } else {
      for (int idx = tid; idx < tb * dr * bb; idx += BT)
        S.R[idx] = rs[(int64_t)(y + 1) * g.ts + idx];
    }
    __syncthreads();
    // X1[l][a'][u][b] = sum_a L[a][l][a'] T[a][u][b]
    for (int idx = tid; idx < dl * ba * du * tb; idx += BT) {
      const int b = idx % tb, u = (idx / tb) % du, ap = (idx / (tb * du)) % ba;
      const int l = idx / (tb * du * ba);
      double v = 0.0;
      for (int a = 0; a < ta; ++a) v = fma(S.L[(a * dl + l) * ba + ap], S.T[(a * du + u) * tb + b], v);
      S.X[idx] = v;
    }
    // Y2[a'][d][b][rr] = sum_b' Bo[a'][d][b'] R[b][rr][b']
    for (int idx = tid; idx < ba * dd * tb * dr; idx += BT) {
      const int rr = idx % dr, b = (idx / dr) % tb, ad = idx / (dr * tb);
      double v = 0.0;
      for (int bp = 0; bp < bb; ++bp) v = fma(S.Bo[ad * bb + bp], S.R[(b * dr + rr) * bb + bp], v);
      S.Y[idx] = v;
    }
    __syncthreads();
    // dS[l][u][d][rr] = sum_{a',b} X1[l][a'][u][b] Y2[a'][d][b][rr] -> G
    {
      const int n = dl * du * dd * dr;
      const int s_ = r * Ly + y;
      const int gbase = cs.doff[s_] + cs.p[s_] * n;
      const int pr = cs.pr[s_];
      const double coef = sg * exp(ll + lry + base);
      for (int idx = tid; idx < n; idx += BT) {
        const int rr = idx % dr, d = (idx / dr) % dd, u = (idx / (dr * dd)) % du;
        const int l = idx / (dr * dd * du);
        const int k = g.dmap[gbase + idx];
        if (k < 0) continue;
        double v = 0.0;
        for (int ap = 0; ap < ba; ++ap)
          for (int b = 0; b < tb; ++b)
            v = fma(S.X[((l * ba + ap) * du + u) * tb + b], S.Y[((ap * dd + d) * tb + b) * dr + rr],
                    v);
        if ((bqpar(d, dd) * (bqpar(rr, dr) + pr)) & 1) v = -v;
        G[w * ldg + k] = coef * v;
      }
    }
    if (y == Ly - 1) break;
    __syncthreads();
    // X2[a'][b][d][rr] = sum_{l,u} X1[l][a'][u][b] S[l][u][d][rr]  (into Y)
    for (int idx = tid; idx < ba * tb * dd * dr; idx += BT) {
      const int rr = idx % dr, d = (idx / dr) % dd, b = (idx / (dr * dd)) % tb;
      const int ap = idx / (dr * dd * tb);
      double v = 0.0;
      for (int l = 0; l < dl; ++l)
        for (int u = 0; u < du; ++u)
          v = fma(S.X[((l * ba + ap) * du + u) * tb + b], S.Sl[((l * du + u) * dd + d) * dr + rr],
                  v);
      S.Y[idx] = v;
    }
    __syncthreads();
    // L_{y+1}[b][rr][b'] = sum_{a',d} X2[a'][b][d][rr] Bo[a'][d][b']
    double mx = 0.0;
    const int nl = tb * dr * bb;
    for (int idx = tid; idx < nl; idx += BT) {
      const int bp = idx % bb, rr = (idx / bb) % dr, b = idx / (bb * dr);
      double v = 0.0;
      for (int ap = 0; ap < ba; ++ap)
        for (int d = 0; d < dd; ++d)
          v = fma(S.Y[((ap * tb + b) * dd + d) * dr + rr], S.Bo[(ap * dd + d) * bb + bp], v);
      S.X[idx] = v;
      mx = fmax(mx, fabs(v));
    }
    mx = bm_max(mx, S.red);
    if (!(mx > 0.0) || !std::isfinite(mx)) return;
    const double inv = 1.0 / mx;
    ll += log(mx);
    for (int idx = tid; idx < nl; idx += BT) S.L[idx] = S.X[idx] * inv;
    __syncthreads();
  }
}

// the boundaries an accepted proposal recomputed (buffers start .. nrows-1
// of each side, and their headers) into its walker's cache
__global__ void __launch_bounds__(256)
    bm_commit_kernel(BmGeo g, int64_t rows, const int32_t *__restrict__ acc, int st_top,
                     int st_bot, const double *__restrict__ ws, const int64_t *__restrict__ hdr,
                     double *cache, int64_t *chdr) {
  const int64_t w = blockIdx.x;
  if (w >= rows || !acc[w]) return;
  const int side = blockIdx.y;
  const bool down = side == 0;
  const int h = g.Lx / 2, nrows = down ? h : g.Lx - h;
  const int start = down ? st_top : st_bot;
  if (start >= nrows) return;
  const int bufd = g.nbuf > 1 ? 1 : 0;
  const int64_t bstride = (int64_t)g.Ly * g.ts;
  const double *src = ws + (w * 2 + side) * g.slot;
  double *dst = cache + (w * 2 + side) * g.slot;
  const int64_t *hs = hdr + (w * 2 + side) * g.nbuf * BHDR;
  int64_t *hc = chdr + (w * 2 + side) * g.nbuf * BHDR;
  for (int b = start; b < nrows; ++b) {
    const int64_t o = b * bufd * bstride;
    for (int64_t i = threadIdx.x; i < bstride; i += 256) dst[o + i] = src[o + i];
    for (int i = threadIdx.x; i < BHDR; i += 256) hc[b * bufd * BHDR + i] = hs[b * bufd * BHDR + i];
  }
}

__global__ void bm_zero_rows_kernel(int64_t rows, int64_t P, double *G, int64_t ldg) {
  const int64_t n = rows * P;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t w = i / P, k = i - w * P;
    G[w * ldg + k] = 0.0;
  }
}

// ------------------------------------------------------------ host side
namespace {
BmGeo bm_geo(int Lx, int Ly, int D, int chi, int nbuf, int half = 0) {
  BmGeo g{};
  g.half = half;
  g.Lx = Lx;
  g.Ly = Ly;
  g.D = D;
  g.N = Lx * Ly;
  g.chi = chi;
  g.cb = chi > D ? chi : D;
  g.nbuf = nbuf;
  g.ts = (int64_t)g.cb * D * g.cb;
  g.slot = (int64_t)nbuf * Ly * g.ts + (int64_t)Ly * BN * BN;
  return g;
}
// the side kernel with 32 x 32 working matrices when chi * D allows (about
// 45 KB of LDS, several workgroups per CU), else 64 x 64 (one per CU)
void launch_side(const BmGeo &g, int64_t n, const int64_t *c, const int32_t *f, double *slots,
                 int64_t *hdr, hipStream_t st, int st_top = 0, int st_bot = 0,
                 const double *cache = nullptr, const int64_t *chdr = nullptr,
                 const int32_t *parent = nullptr, const uint32_t *cmove = nullptr) {
  if (g.cb * g.D <= 32)
    hipLaunchKernelGGL(peps_bmps_side_kernel<32>, dim3((unsigned)n, 2), dim3(BT), 0, st, g, n,
                       c, f, slots, hdr, st_top, st_bot, cache, chdr, parent, cmove);
  else
    hipLaunchKernelGGL(peps_bmps_side_kernel<64>, dim3((unsigned)n, 2), dim3(BT), 0, st, g, n,
                       c, f, slots, hdr, st_top, st_bot, cache, chdr, parent, cmove);
}
}  // namespace

const char *bmps_check(int Lx, int Ly, int D, int chi) {
  if (chi < 1 || chi > BMAX_CHI) return "max_bond (chi) must be 1..32";
  if (Lx < 2 || Lx > 16 || Ly < 1 || Ly > BMAX_LY || Lx * Ly > BMAX_NS)
    return "boundary MPS needs 2 <= Lx <= 16, 1 <= Ly <= 16, Lx * Ly <= 64";
  if (D < 1 || D > 4) return "D must be 1..4";
  const int cb = chi > D ? chi : D;
  if (cb * D > BN || cb * D * cb > BMAX_TS) return "max(chi, D) * D must be <= 64";
  return nullptr;
}

int64_t bmps_workspace_bytes(int Lx, int Ly, int D, int chi, int64_t rows) {
  const BmGeo g = bm_geo(Lx, Ly, D, chi, 1);
  return rows * 2 * (g.slot * 8 + BHDR * 8);
}

int64_t bmps_logderiv_workspace_bytes(int Lx, int Ly, int D, int chi, int64_t rows) {
  const BmGeo g = bm_geo(Lx, Ly, D, chi, Lx - 1);
  return rows * 2 * (g.slot * 8 + g.nbuf * BHDR * 8) +
         rows * Lx * ((int64_t)(Ly + 1) * g.ts + Ly + 1) * 8;
}

int bmps_amplitudes(int Lx, int Ly, int D, int chi, const int32_t *dmap, const double *theta,
                    int64_t theta_stride, int64_t rows, const int64_t *cfg, const int32_t *flag,
                    double *sign, double *logabs, void *ws, int64_t rows_per_pass,
                    hipStream_t st) {
  BmGeo g = bm_geo(Lx, Ly, D, chi, 1);
  g.dmap = dmap;
  double *slots = (double *)ws;
  int64_t *hdr = (int64_t *)(slots + rows_per_pass * 2 * g.slot);
  for (int64_t r0 = 0; r0 < rows; r0 += rows_per_pass) {
    const int64_t n = rows - r0 < rows_per_pass ? rows - r0 : rows_per_pass;
    g.theta = theta + r0 * theta_stride;
    g.theta_stride = theta_stride;
    const int64_t *c = cfg + r0 * g.N;
    const int32_t *f = flag ? flag + r0 : nullptr;
    launch_side(g, n, c, f, slots, hdr, st);
    VMCX_LAUNCHED();
    hipLaunchKernelGGL(peps_bmps_join_kernel, dim3((unsigned)n), dim3(64), 0, st, g, n, c, f,
                       slots, hdr, 0, 0, sign + r0, logabs + r0, 0, 0, nullptr, nullptr, nullptr,
                       nullptr);
    VMCX_LAUNCHED();
  }
  return 0;
}

// ---- the sweep's boundary cache (plain fPEPS: the site tensors do not
// depend on the configuration).  Every walker keeps the boundary of each
// side after each of its rows (top: rows 0 .. h-1, bottom: Lx-1 .. h); a
// proposal on an edge between rows rmin <= rmax recomputes a side only
// from its first changed row on, starting from the walker's boundary before
// it, and takes an unchanged side from the cache as it is.  The arithmetic
// of a recomputed row is that of the full contraction on the same inputs, so
// the amplitudes equal the full recomputation's bit for bit.
namespace {
BmGeo bm_cache_geo(int Lx, int Ly, int D, int chi) {
  const int h = Lx / 2;
  return bm_geo(Lx, Ly, D, chi, h > Lx - h ? h : Lx - h, 1);
}
struct BmCacheWs {
  double *cache, *prop;
  int64_t *chdr, *phdr;
  int32_t *acc;
};
BmCacheWs bm_cache_ws(const BmGeo &g, int64_t B, void *ws) {
  BmCacheWs c;
  const int64_t sd = B * 2 * g.slot;
  const int64_t hw = B * 2 * g.nbuf * BHDR;
  c.cache = (double *)ws;
  c.prop = c.cache + sd;
  c.chdr = (int64_t *)(c.prop + sd);
  c.phdr = c.chdr + hw;
  c.acc = (int32_t *)(c.phdr + hw);
  return c;
}
}  // namespace

int64_t bmps_sweep_cache_bytes(int Lx, int Ly, int D, int chi, int64_t B) {
  const BmGeo g = bm_cache_geo(Lx, Ly, D, chi);
  return 2 * B * 2 * g.slot * 8 + 2 * B * 2 * g.nbuf * BHDR * 8 + B * 4;
}

int32_t *bmps_sweep_accept_flags(int Lx, int Ly, int D, int chi, int64_t B, void *ws) {
  return bm_cache_ws(bm_cache_geo(Lx, Ly, D, chi), B, ws).acc;
}

int bmps_sweep_init(int Lx, int Ly, int D, int chi, const int32_t *dmap, const double *theta,
                    int64_t B, const int64_t *cfg, void *ws, hipStream_t st) {
  BmGeo g = bm_cache_geo(Lx, Ly, D, chi);
  g.dmap = dmap;
  g.theta = theta;
  g.theta_stride = 0;
  const BmCacheWs c = bm_cache_ws(g, B, ws);
  launch_side(g, B, cfg, nullptr, c.cache, c.chdr, st);
  VMCX_LAUNCHED();
  return 0;
}

// the side steps an edge between rows r0 and r1 changes
static void bm_edge_starts(int Lx, int r0, int r1, int &st_top, int &st_bot) {
  const int h = Lx / 2;
  const int rmin = r0 < r1 ? r0 : r1, rmax = r0 < r1 ? r1 : r0;
  st_top = rmin <= h - 1 ? rmin : h;
  st_bot = rmax >= h ? Lx - 1 - rmax : Lx - h;
}

int bmps_sweep_amplitudes(int Lx, int Ly, int D, int chi, const int32_t *dmap,
                          const double *theta, int64_t B, const int64_t *pcfg,
                          const int32_t *flag, int r0, int r1, void *ws, double *sign,
                          double *logabs, hipStream_t st) {
  BmGeo g = bm_cache_geo(Lx, Ly, D, chi);
  g.dmap = dmap;
  g.theta = theta;
  g.theta_stride = 0;
  const BmCacheWs c = bm_cache_ws(g, B, ws);
  int st_top, st_bot;
  bm_edge_starts(Lx, r0, r1, st_top, st_bot);
  const int h = Lx / 2;
  launch_side(g, B, pcfg, flag, c.prop, c.phdr, st, st_top, st_bot, c.cache, c.chdr);
  VMCX_LAUNCHED();
  hipLaunchKernelGGL(peps_bmps_join_kernel, dim3((unsigned)B), dim3(64), 0, st, g, B, pcfg, flag,
                     c.prop, c.phdr, h - 1, Lx - h - 1, sign, logabs, st_top >= h ? 1 : 0,
                     st_bot >= Lx - h ? 1 : 0, c.cache, c.chdr, nullptr, nullptr);
  VMCX_LAUNCHED();
  return 0;
}

// amplitudes of T connected configurations, each one move (cmove) away from
// its walker cparent, from the walkers' cached boundaries (bmps_sweep_init
// of the walkers first), in passes of B through the proposal slots
int bmps_conn_amplitudes(int Lx, int Ly, int D, int chi, const int32_t *dmap,
                         const double *theta, int64_t B, int64_t T, const int64_t *etas,
                         const int32_t *cparent, const uint32_t *cmove, void *ws, double *sign,
                         double *logabs, hipStream_t st) {
  BmGeo g = bm_cache_geo(Lx, Ly, D, chi);
  g.dmap = dmap;
  g.theta = theta;
  g.theta_stride = 0;
  const BmCacheWs c = bm_cache_ws(g, B, ws);
  const int h = Lx / 2;
  for (int64_t r0 = 0; r0 < T; r0 += B) {
    const int64_t n = T - r0 < B ? T - r0 : B;
    const int64_t *e = etas + r0 * g.N;
    launch_side(g, n, e, nullptr, c.prop, c.phdr, st, 0, 0, c.cache, c.chdr, cparent + r0,
                cmove + r0);
    VMCX_LAUNCHED();
    hipLaunchKernelGGL(peps_bmps_join_kernel, dim3((unsigned)n), dim3(64), 0, st, g, n, e,
                       nullptr, c.prop, c.phdr, h - 1, Lx - h - 1, sign + r0, logabs + r0, 0, 0,
                       c.cache, c.chdr, cparent + r0, cmove + r0);
    VMCX_LAUNCHED();
  }
  return 0;
}

int bmps_sweep_commit(int Lx, int Ly, int D, int chi, int64_t B, int r0, int r1, void *ws,
                      hipStream_t st) {
  const BmGeo g = bm_cache_geo(Lx, Ly, D, chi);
  const BmCacheWs c = bm_cache_ws(g, B, ws);
  int st_top, st_bot;
  bm_edge_starts(Lx, r0, r1, st_top, st_bot);
  hipLaunchKernelGGL(bm_commit_kernel, dim3((unsigned)B, 2), dim3(256), 0, st, g, B, c.acc,
                     st_top, st_bot, c.prop, c.phdr, c.cache, c.chdr);
  VMCX_LAUNCHED();
  return 0;
}

int bmps_logderiv(int Lx, int Ly, int D, int chi, const int32_t *dmap, const double *theta,
                  int64_t theta_stride, int64_t n_params, int64_t rows, const int64_t *cfg,
                  double *G, int64_t ldg, double *sign, double *logabs, void *ws,
                  int64_t rows_per_pass, hipStream_t st) {
  BmGeo g = bm_geo(Lx, Ly, D, chi, Lx - 1);
  g.dmap = dmap;
  double *slots = (double *)ws;
  int64_t *hdr = (int64_t *)(slots + rows_per_pass * 2 * g.slot);
  double *renv = (double *)(hdr + rows_per_pass * 2 * g.nbuf * BHDR);
  const int h = Lx / 2;
  for (int64_t r0 = 0; r0 < rows; r0 += rows_per_pass) {
    const int64_t n = rows - r0 < rows_per_pass ? rows - r0 : rows_per_pass;
    g.theta = theta + r0 * theta_stride;
    g.theta_stride = theta_stride;
    const int64_t *c = cfg + r0 * g.N;
    double *Gr = G + r0 * ldg;
    launch_side(g, n, c, nullptr, slots, hdr, st);
    VMCX_LAUNCHED();
    hipLaunchKernelGGL(peps_bmps_join_kernel, dim3((unsigned)n), dim3(64), 0, st, g, n, c,
                       nullptr, slots, hdr, h - 1, Lx - h - 1, sign + r0, logabs + r0, 0, 0,
                       nullptr, nullptr, nullptr, nullptr);
    VMCX_LAUNCHED();
    const int64_t nz = n * n_params;
    const unsigned gz = (unsigned)((nz + 255) / 256 < 16384 ? (nz + 255) / 256 : 16384);
    hipLaunchKernelGGL(bm_zero_rows_kernel, dim3(gz), dim3(256), 0, st, n, n_params, Gr, ldg);
    VMCX_LAUNCHED();
    hipLaunchKernelGGL(peps_bmps_grad_kernel, dim3((unsigned)n, (unsigned)Lx), dim3(BT), 0, st,
                       g, n, c, nullptr, slots, hdr, sign + r0, logabs + r0, renv, Gr, ldg);
    VMCX_LAUNCHED();
  }
  return 0;
}

}  // namespace vmcx
