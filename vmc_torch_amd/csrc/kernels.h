// Internal (C++) entry points of the kernel files, called by capi.hip.
#pragma once

#include "common.h"

typedef struct _rocblas_handle *rocblas_handle;

namespace vmcx {

struct SweepArgs {
  int64_t B;
  int n_sites, n_f, n_edges, n_sweeps;
  const double *M;
  const int32_t *edges;
  double keep_rate;  // 1 - hopping_rate
  uint64_t seed;
  int64_t sweep_offset;
  const double *u_choice;
  const uint8_t *bits;
  const double *u_acc;
  int64_t *configs;
  double *ainv;
  uint8_t *slot;
  double *sign, *logabs;
  int32_t *since_refresh;
  int64_t *n_accept;
  // the FP32 form: M rounded to FP32 and the FP32 A^-1 (M above stays the
  // FP64 matrix of the exact refresh)
  const float *M32 = nullptr;
  float *ainv32 = nullptr;
  // a device-resident addition to sweep_offset (NULL: none), read by the
  // kernel: lets a captured graph advance the draws between replays
  const int64_t *offset_dev = nullptr;
};

int slater_refresh(int64_t B, int n_sites, int n_f, const double *M,
                   const int64_t *configs, double *ainv, uint8_t *slot,
                   double *sign, double *logabs, int32_t *since,
                   int only_if_since, hipStream_t st);
int slater_sweep(const SweepArgs &a, int refresh_every, hipStream_t st);
int slater_local_energy(int64_t B, int n_sites, int n_f, const double *M,
                        int n_edges, const int32_t *edges, double t, double U,
                        const int64_t *configs, const double *ainv,
                        const uint8_t *slot, double *eloc, hipStream_t st);
int slater_logderiv(int64_t B, int n_sites, int n_f, const double *ainv,
                    const uint8_t *slot, void *O, int o_dtype, int64_t ld,
                    hipStream_t st);
int slater_refresh_f32(int64_t B, int n_sites, int n_f, const double *M,
                       const int64_t *configs, float *ainv, uint8_t *slot, double *sign,
                       double *logabs, int32_t *since, int only_if_since, hipStream_t st);
int slater_sweep_f32(const SweepArgs &a, int refresh_every, hipStream_t st);
int slater_local_energy_f32(int64_t B, int n_sites, int n_f, const float *M, int n_edges,
                            const int32_t *edges, double t, double U, const int64_t *configs,
                            const float *ainv, const uint8_t *slot, double *eloc,
                            hipStream_t st);
int slater_logderiv_f32(int64_t B, int n_sites, int n_f, const float *ainv,
                        const uint8_t *slot, void *O, int o_dtype, int64_t ld,
                        hipStream_t st);
int sgd_update_f32(int64_t n, float *params, const double *dp, double lr, hipStream_t st);
int hubbard_conn_count(int64_t B, int n_sites, int n_edges,
                       const int32_t *edges, const int64_t *configs,
                       int64_t *counts, hipStream_t st);
int hubbard_conn_fill(int64_t B, int n_sites, int n_edges, const int32_t *edges,
                      double t, double U, const int64_t *configs,
                      const int64_t *offsets, int64_t *etas, double *coeffs,
                      int64_t *batch_ids, hipStream_t st);

int64_t sr_workspace_bytes(int64_t Ns, int64_t Np, int o_dtype, int gram, int flags);
int sr_minsr(int64_t Ns, int64_t Np, void *O, int o_dtype, int64_t ldo, int flags,
             const double *E, double energy_mean, double shift, int gram, int max_iter,
             double tol, void *ws, int64_t ws_bytes, double *dp, double *stats, int32_t *info,
             float *phase_ms, hipStream_t st);
int sgd_update(int64_t n, double *params, const double *dp, double lr,
               hipStream_t st);
rocblas_handle blas_handle_for(hipStream_t st);
int64_t sr_minres_workspace_bytes(int64_t Ns, int64_t Np);
int sr_minres(int64_t Ns, int64_t Np, const void *O, int o_dtype, int64_t ldo, const double *E,
              double energy_mean, double diag_shift, double rtol, int maxiter, void *ws,
              int64_t ws_bytes, double *dp, double *stats, int32_t *info, hipStream_t st);
// splits >= 2: split 0 writes C, split s >= 1 writes Crest + (s - 1) cstride
// (Crest == NULL: C + cstride, i.e. contiguous splits)
// yield != NULL: a device word holding the CU code (cu_code below) of a running
// diagonal factorisation kernel, or 0; waves of this SYRK on that CU sleep
// while it runs (at most ~0.15 ms per chunk), leaving it the matrix pipes
int syrk_f64(int64_t n, int64_t K, const double *X, int64_t ldx, int rowmajor, double alpha,
             double beta, double *C, int64_t ldc, int splits, int64_t cstride, hipStream_t st,
             double *Crest = nullptr, const int32_t *yield = nullptr);
#ifdef __HIPCC__
// the calling wave's CU, never 0: XCC id (HW_REG_XCC_ID bits 0-3) and the
// CU / SH / SE fields of HW_REG_HW_ID (bits 8-15)
__device__ __forceinline__ int32_t cu_code() {
  const int hw = __builtin_amdgcn_s_getreg((7 << 11) | (8 << 6) | 4);
  const int xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);
  return 1 + (((xcc & 15) << 8) | (hw & 255));
}
#endif
int64_t chol_dinv_bytes(int64_t n);
// the off-diagonal blocks of the factor (n x n, ld n) inside the dinv workspace
const double *chol_panels(int64_t n, const double *dinv);
// diag_ms != NULL: mean duration (ms) of one diagonal-block launch, timed with
// events on `st` (the call then waits for the stream)
// fwd_b != NULL: also y = L^-1 fwd_b into fwd_y (fwd_b overwritten), computed
// on a third stream while the factorisation proceeds
int chol_factor(void *rocblas_handle_, int64_t n, double *A, int64_t lda,
                double *dinv, int32_t *info, hipStream_t st, float *diag_ms = nullptr,
                double *fwd_b = nullptr, double *fwd_y = nullptr);
// halves of chol_solve: y = L^-1 b (b overwritten), x = L^-T y (y overwritten)
int chol_solve_fwd(int64_t n, const double *L, int64_t lda, const double *dinv, double *b,
                   double *y, hipStream_t st, const double *skip = nullptr);
int chol_solve_bwd(int64_t n, const double *L, int64_t lda, const double *dinv, double *y,
                   double *x, hipStream_t st, const double *skip = nullptr);
// skip != NULL and *skip != 0 (a device flag): every launch returns at once
int chol_solve(int64_t n, const double *L, int64_t lda, const double *dinv,
               const double *b, double *x, double *tmp, double *tmp2,
               hipStream_t st, const double *skip = nullptr);

}  // namespace vmcx

#include "../../include/vmcx.h"

namespace vmcx {
// kernels of nnbf.hip shared by the fPEPS walkers
__global__ void nnbf_propose_kernel(int64_t B, int N, int i, int j, int e, int64_t sweep,
                                    uint64_t seed, double keep_rate, const double *u_choice,
                                    const uint8_t *bits, const double *u_acc_in,
                                    const int64_t *cfg, int32_t *prop, uint32_t *moves,
                                    double *u_acc);
__global__ void conn_count_walker_kernel(int64_t B, int N, int E, const int32_t *edges,
                                         const int64_t *cfg, int64_t *counts);
// cparent / cmove (may be NULL): for connected configuration k, its walker
// and the hop as from_orbital | to_orbital << 8 (orbital = site for up, N +
// site for down)
int launch_conn_fill(int64_t B, int N, int E, const int32_t *edges, double t,
                     const int64_t *cfg, const int64_t *offsets, int64_t *etas, double *coeffs,
                     hipStream_t st, int32_t *cparent = nullptr, uint32_t *cmove = nullptr);
__global__ void exclusive_scan_kernel(int64_t n, const int64_t *in, int64_t *out,
                                      int64_t *total);
__global__ void conn_fill_walker_kernel(int64_t B, int N, int E, const int32_t *edges,
                                        double t, const int64_t *cfg, const int64_t *offsets,
                                        int64_t *etas, double *coeffs, int32_t *cparent,
                                        uint32_t *cmove);
__global__ void eloc_segment_kernel(int64_t B, int N, double U, const int64_t *cfg,
                                    const int64_t *offsets, const int64_t *counts,
                                    const double *coeffs, const double *csign,
                                    const double *clog, const double *sign,
                                    const double *logabs, double *eloc);

int64_t peps_dense_size(const vmcx_peps_model *m);
int64_t peps_workspace_bytes(const vmcx_peps_model *m, int64_t B, int64_t chunk, int n_edges);
int64_t peps_reuse_workspace_bytes(const vmcx_peps_model *m, int64_t B, int64_t chunk,
                                   int n_edges);
int peps_sweep_reuse(const vmcx_peps_model *m, int64_t B, int n_edges, const int32_t *edges_host,
                     double hopping_rate, int n_sweeps, uint64_t seed, int64_t sweep_offset,
                     const double *u_choice, const uint8_t *bits, const double *u_acc,
                     int64_t *cfg, double *sign, double *logabs, int64_t *n_accept, void *ws,
                     int64_t ws_bytes, hipStream_t st);
int peps_local_energy_reuse(const vmcx_peps_model *m, int64_t B, int n_edges,
                            const int32_t *edges, double t, double U, const int64_t *cfg,
                            const double *sign, const double *logabs, void *ws,
                            int64_t ws_bytes, int64_t chunk, double *eloc, hipStream_t st);
int peps_forward(const vmcx_peps_model *m, int64_t B, const int64_t *cfg, void *ws,
                 int64_t ws_bytes, int64_t chunk, double *sign, double *logabs, hipStream_t st);
int peps_sweep(const vmcx_peps_model *m, int64_t B, int n_edges, const int32_t *edges_host,
               double hopping_rate, int n_sweeps, uint64_t seed, int64_t sweep_offset,
               const double *u_choice, const uint8_t *bits, const double *u_acc, int64_t *cfg,
               double *sign, double *logabs, int64_t *n_accept, void *ws, int64_t ws_bytes,
               hipStream_t st);
int peps_local_energy(const vmcx_peps_model *m, int64_t B, int n_edges, const int32_t *edges,
                      double t, double U, const int64_t *cfg, const double *sign,
                      const double *logabs, void *ws, int64_t ws_bytes, int64_t chunk,
                      double *eloc, hipStream_t st);
int peps_logderiv(const vmcx_peps_model *m, int64_t B, const int64_t *cfg, void *ws,
                  int64_t ws_bytes, int64_t chunk, void *O, int o_dtype, int64_t ld,
                  double *sign, double *logabs, hipStream_t st);
int peps_logderiv_factors(const vmcx_peps_model *m, int64_t B, const int64_t *cfg, void *ws,
                          int64_t ws_bytes, int64_t chunk, double *F, int64_t ld,
                          double *sign, double *logabs, hipStream_t st);

// chi-truncated boundary MPS (bmps.hip): nullptr or the reason the shape is
// not supported; workspace for rows_per_pass configurations; amplitudes of
// `rows` configurations in passes of rows_per_pass
const char *bmps_check(int Lx, int Ly, int D, int chi);
int64_t bmps_workspace_bytes(int Lx, int Ly, int D, int chi, int64_t rows_per_pass);
int bmps_amplitudes(int Lx, int Ly, int D, int chi, const int32_t *dmap, const double *theta,
                    int64_t theta_stride, int64_t rows, const int64_t *cfg, const int32_t *flag,
                    double *sign, double *logabs, void *ws, int64_t rows_per_pass,
                    hipStream_t st);
// the sweep's boundary cache of plain fPEPS (B walkers): workspace, the
// accept flags inside it, the walkers' boundaries, one edge's proposals
// (rows r0, r1 changed; amplitudes bit-identical to the full contraction)
// and the accepted proposals' boundaries into the cache
int64_t bmps_sweep_cache_bytes(int Lx, int Ly, int D, int chi, int64_t B);
int32_t *bmps_sweep_accept_flags(int Lx, int Ly, int D, int chi, int64_t B, void *ws);
int bmps_sweep_init(int Lx, int Ly, int D, int chi, const int32_t *dmap, const double *theta,
                    int64_t B, const int64_t *cfg, void *ws, hipStream_t st);
int bmps_sweep_amplitudes(int Lx, int Ly, int D, int chi, const int32_t *dmap,
                          const double *theta, int64_t B, const int64_t *pcfg,
                          const int32_t *flag, int r0, int r1, void *ws, double *sign,
                          double *logabs, hipStream_t st);
int bmps_sweep_commit(int Lx, int Ly, int D, int chi, int64_t B, int r0, int r1, void *ws,
                      hipStream_t st);
int bmps_conn_amplitudes(int Lx, int Ly, int D, int chi, const int32_t *dmap,
                         const double *theta, int64_t B, int64_t T, const int64_t *etas,
                         const int32_t *cparent, const uint32_t *cmove, void *ws, double *sign,
                         double *logabs, hipStream_t st);
// d ln|psi| / d theta with the boundary environments held fixed (G rows of
// n_params at stride ldg, FP64), and (sign, logabs)
int64_t bmps_logderiv_workspace_bytes(int Lx, int Ly, int D, int chi, int64_t rows_per_pass);
int bmps_logderiv(int Lx, int Ly, int D, int chi, const int32_t *dmap, const double *theta,
                  int64_t theta_stride, int64_t n_params, int64_t rows, const int64_t *cfg,
                  double *G, int64_t ldg, double *sign, double *logabs, void *ws,
                  int64_t rows_per_pass, hipStream_t st);

int lp_accept_commit(int64_t B, int N, uint64_t seed, int64_t step, const double *u_acc,
                     const int32_t *move, const int64_t *pcfg, const double *psign,
                     const double *plog, const double *logq_f, const double *logq_r,
                     int64_t *cfg, double *sign, double *logabs, int64_t *n_accept,
                     uint8_t *accepted, const int64_t *step_dev, int M, const double *logits_p,
                     double *logits_x, int64_t row_bytes, const void *z1p, void *z1,
                     hipStream_t st);
int lp_n_params(int C);
int lp_n_moves(int Lx, int Ly);
int lp_logits(const float *P, int C, int Lx, int Ly, int64_t B, const int64_t *cfg,
              double *logits, int precision, hipStream_t st);
int lp_propose(const float *P, int C, int Lx, int Ly, int64_t B, const int64_t *cfg,
               uint64_t seed, int64_t step, const double *u_choice, double *logits,
               int32_t *move, double *logq_f, int64_t *pcfg, double *logq_r, int precision,
               hipStream_t st);
int lp_accept(int64_t B, int N, uint64_t seed, int64_t step, const double *u_acc,
              const int32_t *move, const int64_t *pcfg, const double *psign,
              const double *plog, const double *logq_f, const double *logq_r, int64_t *cfg,
              double *sign, double *logabs, int64_t *n_accept, uint8_t *accepted,
              hipStream_t st, const int64_t *step_dev = nullptr);
int counter_add(int64_t *ctr, int64_t inc, hipStream_t st);
int64_t lp_backward_workspace_bytes(int C, int64_t B);
int lp_backward(const float *P, int C, int Lx, int Ly, int64_t B, const int64_t *cfg,
                const double *dlogits, void *ws, int64_t ws_bytes, float *grad, hipStream_t st);
int lp_candidates(int64_t B, int Lx, int Ly, const int64_t *cfg, int64_t *out, hipStream_t st);
int lp_axpy(int64_t n, float a, const float *x, float *y, hipStream_t st);

int lp_propose_given(const float *P, int C, int Lx, int Ly, int64_t B, const int64_t *cfg,
                     const double *logits_x, uint64_t seed, int64_t step,
                     const double *u_choice, double *logits_p, int32_t *move, double *logq_f,
                     int64_t *pcfg, double *logq_r, int precision, hipStream_t st,
                     const int64_t *step_dev = nullptr);
int lp_commit_logits(int64_t B, int M, const uint8_t *accepted, const double *logits_p,
                     double *logits_x, hipStream_t st);

int sym_images(int64_t B, int N, int G, const int64_t *cfg, const int32_t *perm,
               const int32_t *inv, int64_t *images, double *sign, hipStream_t st);

int64_t nnbf_workspace_bytes(const vmcx_nnbf_model *m, int64_t B, int64_t chunk, int n_edges);
int nnbf_forward_moved(const vmcx_nnbf_model *m, int64_t B, const int64_t *cfg,
                       const int64_t *pcfg, const void *z1, void *ws, int64_t ws_bytes,
                       int64_t chunk, double *sign, double *logabs, void *z1p, hipStream_t st);
int masked_rows(int64_t B, int64_t row_bytes, const uint8_t *mask, const void *src, void *dst,
                hipStream_t st);
int nnbf_forward(const vmcx_nnbf_model *m, int64_t B, const int64_t *cfg, void *ws,
                 int64_t ws_bytes, int64_t chunk, double *sign, double *logabs, void *z1_out,
                 hipStream_t st);
int nnbf_sweep(const vmcx_nnbf_model *m, int64_t B, int n_edges, const int32_t *edges_host,
               double hopping_rate, int n_sweeps, uint64_t seed, int64_t sweep_offset,
               const double *u_choice, const uint8_t *bits, const double *u_acc,
               int64_t *cfg, void *z1, double *sign, double *logabs, int64_t *n_accept,
               void *ws, int64_t ws_bytes, hipStream_t st);
int nnbf_local_energy(const vmcx_nnbf_model *m, int64_t B, int n_edges, const int32_t *edges,
                      double t, double U, const int64_t *cfg, const double *sign,
                      const double *logabs, void *ws, int64_t ws_bytes, int64_t chunk,
                      double *eloc, hipStream_t st);
int nnbf_logderiv(const vmcx_nnbf_model *m, int64_t B, const int64_t *cfg, void *ws,
                  int64_t ws_bytes, int64_t chunk, void *O, int o_dtype, int64_t ld,
                  double *sign, double *logabs, hipStream_t st);
int nnbf_logderiv_factors(const vmcx_nnbf_model *m, int64_t B, const int64_t *cfg, void *ws,
                          int64_t ws_bytes, int64_t chunk, double *F, int64_t ld,
                          double *sign, double *logabs, hipStream_t st);
int to_bf16(int64_t n, const float *x, void *y, hipStream_t st);
int64_t sr_factored_workspace_bytes(int64_t Ns, int n_factors, const int32_t *k, int n_blocks,
                                    const vmcx_sr_block *blocks);
int sr_minsr_factored(int64_t Ns, int n_factors, const double *const *F, const int64_t *ld,
                      const int32_t *k, int n_blocks, const vmcx_sr_block *blocks, int64_t Np,
                      const double *E, double energy_mean, double shift, void *ws,
                      int64_t ws_bytes, double *dp, double *stats, int32_t *info,
                      float *phase_ms, hipStream_t st);
int sr_factored_gradient(int64_t Ns, int n_factors, const double *const *F, const int64_t *ld,
                         const int32_t *k, int n_blocks, const vmcx_sr_block *blocks, int64_t Np,
                         const double *E, double energy_mean, void *ws, int64_t ws_bytes,
                         double *g, double *stats, hipStream_t st);
int energy_stats(int64_t n, const double *E, double *out, hipStream_t st);
}  // namespace vmcx
