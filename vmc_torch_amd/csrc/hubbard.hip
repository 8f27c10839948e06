// Batched connected configurations of the spinful Hubbard Hamiltonian in
// the reference's term-major order (hamiltonian_torch.py:586-655):
//   term 2e   : up hop on edge e     term 2e+1 : down hop on edge e
//   term 2E+s : on-site U at site s (only where the site is doubly occupied)
// Within a term, walkers appear in ascending order.
#include "kernels.h"

namespace vmcx {

__device__ __forceinline__ bool term_valid(const int64_t *cfg, int term,
                                           int n_sites, int n_edges,
                                           const int32_t *edges) {
  if (term < 2 * n_edges) {
    const int e = term >> 1, bit = (term & 1) ? 1 : 2;  // down : up
    const int i = edges[2 * e], j = edges[2 * e + 1];
    if ((unsigned)i >= (unsigned)n_sites || (unsigned)j >= (unsigned)n_sites || i == j)
      return false;
    return ((cfg[i] & bit) != 0) != ((cfg[j] & bit) != 0);
  }
  return cfg[term - 2 * n_edges] == 3;
}

__global__ void zero_i64(int64_t *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// grid: (n_terms, ceil(B/256)); counts[term] += #valid walkers
__global__ void __launch_bounds__(256)
    conn_count_kernel(int64_t B, int n_sites, int n_edges,
                      const int32_t *__restrict__ edges,
                      const int64_t *__restrict__ configs, int64_t *counts) {
  const int term = blockIdx.x;
  const int64_t b = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
  const bool v = (b < B) && term_valid(configs + b * n_sites, term, n_sites, n_edges, edges);
  const uint64_t bal = __ballot(v);
  __shared__ int part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = __popcll(bal);
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += part[k];
    if (s) atomicAdd((unsigned long long *)&counts[term], (unsigned long long)s);
  }
}

// one block (1024 threads) per term; block-wide ordered compaction
__global__ void __launch_bounds__(1024)
    conn_fill_kernel(int64_t B, int n_sites, int n_edges,
                     const int32_t *__restrict__ edges, double t, double U,
                     const int64_t *__restrict__ configs,
                     const int64_t *__restrict__ offsets, int64_t *etas,
                     double *coeffs, int64_t *batch_ids) {
  const int term = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  __shared__ int wcount[16];
  int64_t base = offsets[term];
  const bool hop = term < 2 * n_edges;
  int i = 0, j = 0, bit = 0;
  if (hop) {
    const int e = term >> 1;
    bit = (term & 1) ? 1 : 2;
    i = edges[2 * e];
    j = edges[2 * e + 1];
  }
  for (int64_t c0 = 0; c0 < B; c0 += blockDim.x) {
    const int64_t b = c0 + threadIdx.x;
    const int64_t *cfg = configs + (b < B ? b : 0) * n_sites;
    const bool v = (b < B) && term_valid(cfg, term, n_sites, n_edges, edges);
    const uint64_t bal = __ballot(v);
    if (lane == 0) wcount[wave] = __popcll(bal);
    __syncthreads();
    int before = 0, total = 0;
    for (int k = 0; k < nw; ++k) {
      const int c = wcount[k];
      if (k < wave) before += c;
      total += c;
    }
    const int rank = before + __popcll(bal & ((1ull << lane) - 1ull));
    if (v) {
      const int64_t pos = base + rank;
      int64_t *eta = etas + pos * n_sites;
      for (int s = 0; s < n_sites; ++s) eta[s] = cfg[s];
      double coef;
      if (hop) {
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        // symmray modes strictly between the two hopping modes
        int par = 0;
        if (bit == 2) {  // up: ups in (lo, hi), downs in (lo, hi]
          for (int s = lo + 1; s < hi; ++s) par += (int)((cfg[s] >> 1) & 1);
          for (int s = lo + 1; s <= hi; ++s) par += (int)(cfg[s] & 1);
        } else {         // down: ups in [lo, hi), downs in (lo, hi)
          for (int s = lo; s < hi; ++s) par += (int)((cfg[s] >> 1) & 1);
          for (int s = lo + 1; s < hi; ++s) par += (int)(cfg[s] & 1);
        }
        coef = (par & 1) ? t : -t;
        eta[i] = cfg[i] ^ bit;
        eta[j] = cfg[j] ^ bit;
      } else {
        coef = U;
      }
      coeffs[pos] = coef;
      batch_ids[pos] = b;
    }
    base += total;
    __syncthreads();
  }
}

int hubbard_conn_count(int64_t B, int n_sites, int n_edges,
                       const int32_t *edges, const int64_t *configs,
                       int64_t *counts, hipStream_t st) {
  VMCX_CHECK_ARG(B >= 0 && n_edges >= 0, "negative size");
  VMCX_CHECK_ARG(n_sites >= 1 && n_sites <= 64, "n_sites must be in [1, 64]");
  VMCX_CHECK_ARG(counts && (B == 0 || configs) && (n_edges == 0 || edges),
                 "null pointer");
  const int n_terms = 2 * n_edges + n_sites;
  hipLaunchKernelGGL(zero_i64, dim3((n_terms + 255) / 256), dim3(256), 0, st,
                     counts, n_terms);
  VMCX_LAUNCHED();
  if (B == 0) return 0;
  const int64_t nby = (B + 255) / 256;
  VMCX_CHECK_ARG(nby <= 65535, "B too large for the count grid");
  hipLaunchKernelGGL(conn_count_kernel, dim3(n_terms, (unsigned)nby), dim3(256),
                     0, st, B, n_sites, n_edges, edges, configs, counts);
  VMCX_LAUNCHED();
  return 0;
}

int hubbard_conn_fill(int64_t B, int n_sites, int n_edges, const int32_t *edges,
                      double t, double U, const int64_t *configs,
                      const int64_t *offsets, int64_t *etas, double *coeffs,
                      int64_t *batch_ids, hipStream_t st) {
  VMCX_CHECK_ARG(B >= 0 && n_edges >= 0, "negative size");
  VMCX_CHECK_ARG(n_sites >= 1 && n_sites <= 64, "n_sites must be in [1, 64]");
  VMCX_CHECK_ARG(offsets && (n_edges == 0 || edges), "null pointer");
  if (B == 0) return 0;
  VMCX_CHECK_ARG(configs && etas && coeffs && batch_ids, "null pointer");
  const int n_terms = 2 * n_edges + n_sites;
  hipLaunchKernelGGL(conn_fill_kernel, dim3(n_terms), dim3(1024), 0, st, B,
                     n_sites, n_edges, edges, t, U, configs, offsets, etas,
                     coeffs, batch_ids);
  VMCX_LAUNCHED();
  return 0;
}

}  // namespace vmcx
