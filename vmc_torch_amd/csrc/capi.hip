// extern "C" entry points of libvmcx.so (declared in include/vmcx.h).
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vmcx.h"
#include "kernels.h"

namespace vmcx {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

size_t max_lds_per_block() {
  static std::mutex mu;
  static std::vector<size_t> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  std::lock_guard<std::mutex> lk(mu);
  if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
  if (cache[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock,
                              dev) != hipSuccess || v <= 0)
      v = 65536;
    cache[dev] = (size_t)v;
  }
  return cache[dev];
}

// launch-shape table: (kind, key) -> waves per workgroup (1, 2, 4 or 8)
static std::mutex g_tune_mu;
static std::vector<std::vector<int>> g_tune(TUNE_KINDS);

int tuned_waves(int kind, int key, int dflt) {
  if (kind < 0 || kind >= TUNE_KINDS || key < 0) return dflt;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const std::vector<int> &t = g_tune[kind];
  return (key < (int)t.size() && t[key] > 0) ? t[key] : dflt;
}

int set_tuned_waves(int kind, int key, int waves) {
  VMCX_CHECK_ARG(kind >= 0 && kind < TUNE_KINDS, "unknown kernel kind");
  VMCX_CHECK_ARG(key >= 0 && key < 1024, "key out of range");
  VMCX_CHECK_ARG(waves == 0 || waves == 1 || waves == 2 || waves == 4 || waves == 8,
                 "waves per workgroup must be 0 (default), 1, 2, 4 or 8");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  std::vector<int> &t = g_tune[kind];
  if ((int)t.size() <= key) t.resize(key + 1, 0);
  t[key] = waves;
  return 0;
}

}  // namespace vmcx

static inline int status_of(int rc) {
  return rc == 0 ? VMCX_OK : (rc == 1 ? VMCX_EINVAL : VMCX_EHIP);
}

extern "C" {

int vmcx_abi_version(void) { return 19; }

const char *vmcx_last_error(void) { return vmcx::g_last_error.c_str(); }

int vmcx_slater_nfp(int n_f) { return (n_f >= 1) ? vmcx::nfp_of(n_f) : -1; }

int vmcx_set_launch_waves(int kind, int key, int waves) {
  return status_of(vmcx::set_tuned_waves(kind, key, waves));
}

int vmcx_get_launch_waves(int kind, int key) { return vmcx::tuned_waves(kind, key, 0); }

int vmcx_slater_refresh(int64_t B, int n_sites, int n_f, const double *M,
                        const int64_t *configs, double *ainv, uint8_t *slot,
                        double *sign, double *logabs, int32_t *since_refresh,
                        int only_if_since, void *stream) {
  return status_of(vmcx::slater_refresh(B, n_sites, n_f, M, configs, ainv, slot,
                                        sign, logabs, since_refresh,
                                        only_if_since, (hipStream_t)stream));
}

int vmcx_slater_sweep(int64_t B, int n_sites, int n_f, const double *M,
                      int n_edges, const int32_t *edges, double hopping_rate,
                      int n_sweeps, uint64_t seed, int64_t sweep_offset,
                      const double *u_choice, const uint8_t *bits,
                      const double *u_acc, int refresh_every, int64_t *configs,
                      double *ainv, uint8_t *slot, double *sign, double *logabs,
                      int32_t *since_refresh, int64_t *n_accept,
                      const int64_t *sweep_offset_dev, void *stream) {
  vmcx::SweepArgs a;
  a.offset_dev = sweep_offset_dev;
  a.B = B;
  a.n_sites = n_sites;
  a.n_f = n_f;
  a.n_edges = n_edges;
  a.n_sweeps = n_sweeps;
  a.M = M;
  a.edges = edges;
  a.keep_rate = 1.0 - hopping_rate;
  a.seed = seed;
  a.sweep_offset = sweep_offset;
  a.u_choice = u_choice;
  a.bits = bits;
  a.u_acc = u_acc;
  a.configs = configs;
  a.ainv = ainv;
  a.slot = slot;
  a.sign = sign;
  a.logabs = logabs;
  a.since_refresh = since_refresh;
  a.n_accept = n_accept;
  return status_of(vmcx::slater_sweep(a, refresh_every, (hipStream_t)stream));
}

int vmcx_slater_refresh_f32(int64_t B, int n_sites, int n_f, const double *M,
                            const int64_t *configs, float *ainv, uint8_t *slot, double *sign,
                            double *logabs, int32_t *since_refresh, int only_if_since,
                            void *stream) {
  return status_of(vmcx::slater_refresh_f32(B, n_sites, n_f, M, configs, ainv, slot, sign,
                                            logabs, since_refresh, only_if_since,
                                            (hipStream_t)stream));
}

int vmcx_slater_sweep_f32(int64_t B, int n_sites, int n_f, const double *M, const float *M32,
                          int n_edges, const int32_t *edges, double hopping_rate, int n_sweeps,
                          uint64_t seed, int64_t sweep_offset, const double *u_choice,
                          const uint8_t *bits, const double *u_acc, int refresh_every,
                          int64_t *configs, float *ainv, uint8_t *slot, double *sign,
                          double *logabs, int32_t *since_refresh, int64_t *n_accept,
                          const int64_t *sweep_offset_dev, void *stream) {
  vmcx::SweepArgs a;
  a.offset_dev = sweep_offset_dev;
  a.B = B;
  a.n_sites = n_sites;
  a.n_f = n_f;
  a.n_edges = n_edges;
  a.n_sweeps = n_sweeps;
  a.M = M;
  a.M32 = M32;
  a.edges = edges;
  a.keep_rate = 1.0 - hopping_rate;
  a.seed = seed;
  a.sweep_offset = sweep_offset;
  a.u_choice = u_choice;
  a.bits = bits;
  a.u_acc = u_acc;
  a.configs = configs;
  a.ainv = nullptr;
  a.ainv32 = ainv;
  a.slot = slot;
  a.sign = sign;
  a.logabs = logabs;
  a.since_refresh = since_refresh;
  a.n_accept = n_accept;
  return status_of(vmcx::slater_sweep_f32(a, refresh_every, (hipStream_t)stream));
}

int vmcx_slater_local_energy_f32(int64_t B, int n_sites, int n_f, const float *M32,
                                 int n_edges, const int32_t *edges, double t, double U,
                                 const int64_t *configs, const float *ainv,
                                 const uint8_t *slot, double *eloc, void *stream) {
  return status_of(vmcx::slater_local_energy_f32(B, n_sites, n_f, M32, n_edges, edges, t, U,
                                                 configs, ainv, slot, eloc,
                                                 (hipStream_t)stream));
}

int vmcx_slater_logderiv_f32(int64_t B, int n_sites, int n_f, const float *ainv,
                             const uint8_t *slot, void *O, int o_dtype, int64_t ld,
                             void *stream) {
  return status_of(vmcx::slater_logderiv_f32(B, n_sites, n_f, ainv, slot, O, o_dtype, ld,
                                             (hipStream_t)stream));
}

int vmcx_sgd_update_f32(int64_t n, float *params, const double *dp, double lr, void *stream) {
  return status_of(vmcx::sgd_update_f32(n, params, dp, lr, (hipStream_t)stream));
}

int vmcx_slater_local_energy(int64_t B, int n_sites, int n_f, const double *M,
                             int n_edges, const int32_t *edges, double t,
                             double U, const int64_t *configs,
                             const double *ainv, const uint8_t *slot,
                             double *eloc, void *stream) {
  return status_of(vmcx::slater_local_energy(B, n_sites, n_f, M, n_edges, edges,
                                             t, U, configs, ainv, slot, eloc,
                                             (hipStream_t)stream));
}

int vmcx_slater_logderiv(int64_t B, int n_sites, int n_f, const double *ainv,
                         const uint8_t *slot, void *O, int o_dtype, int64_t ld,
                         void *stream) {
  return status_of(vmcx::slater_logderiv(B, n_sites, n_f, ainv, slot, O,
                                         o_dtype, ld, (hipStream_t)stream));
}

int vmcx_hubbard_conn_count(int64_t B, int n_sites, int n_edges,
                            const int32_t *edges, const int64_t *configs,
                            int64_t *counts, void *stream) {
  return status_of(vmcx::hubbard_conn_count(B, n_sites, n_edges, edges, configs,
                                            counts, (hipStream_t)stream));
}

int vmcx_hubbard_conn_fill(int64_t B, int n_sites, int n_edges,
                           const int32_t *edges, double t, double U,
                           const int64_t *configs, const int64_t *offsets,
                           int64_t *etas, double *coeffs, int64_t *batch_ids,
                           void *stream) {
  return status_of(vmcx::hubbard_conn_fill(B, n_sites, n_edges, edges, t, U,
                                           configs, offsets, etas, coeffs,
                                           batch_ids, (hipStream_t)stream));
}

int64_t vmcx_sr_workspace_bytes(int64_t Ns, int64_t Np, int o_dtype, int gram_dtype,
                                int flags) {
  return vmcx::sr_workspace_bytes(Ns, Np, o_dtype, gram_dtype, flags);
}

int vmcx_sr_minsr(int64_t Ns, int64_t Np, void *O, int o_dtype, int64_t ldo, int flags,
                  const double *E, double energy_mean, double diag_shift, int gram_dtype,
                  int max_iter, double tol, void *workspace, int64_t ws_bytes, double *dp,
                  double *stats, int32_t *info, float *phase_ms, void *stream) {
  const int rc = vmcx::sr_minsr(Ns, Np, O, o_dtype, ldo, flags, E, energy_mean, diag_shift,
                                gram_dtype, max_iter, tol, workspace, ws_bytes, dp, stats, info,
                                phase_ms, (hipStream_t)stream);
  return rc == 3 ? VMCX_ELIB : status_of(rc);
}

int vmcx_syrk_f64(int64_t n, int64_t K, const double *X, int64_t ldx, int rowmajor,
                  double alpha, double beta, double *C, int64_t ldc, int splits,
                  int64_t cstride, void *stream) {
  return status_of(vmcx::syrk_f64(n, K, X, ldx, rowmajor, alpha, beta, C, ldc, splits, cstride,
                                  (hipStream_t)stream));
}

int64_t vmcx_sr_minres_workspace_bytes(int64_t Ns, int64_t Np) {
  return vmcx::sr_minres_workspace_bytes(Ns, Np);
}

int vmcx_sr_minres(int64_t Ns, int64_t Np, const void *O, int o_dtype, int64_t ldo,
                   const double *E, double energy_mean, double diag_shift, double rtol,
                   int maxiter, void *workspace, int64_t ws_bytes, double *dp, double *stats,
                   int32_t *info, void *stream) {
  return status_of(vmcx::sr_minres(Ns, Np, O, o_dtype, ldo, E, energy_mean, diag_shift, rtol,
                                   maxiter, workspace, ws_bytes, dp, stats, info,
                                   (hipStream_t)stream));
}

int64_t vmcx_nnbf_workspace_bytes(const vmcx_nnbf_model *m, int64_t B, int64_t chunk,
                                  int n_edges) {
  return vmcx::nnbf_workspace_bytes(m, B, chunk, n_edges);
}

static inline int status3(int rc) { return rc == 3 ? VMCX_ELIB : status_of(rc); }

int vmcx_nnbf_forward(const vmcx_nnbf_model *m, int64_t B, const int64_t *configs,
                      void *workspace, int64_t ws_bytes, int64_t chunk, double *sign,
                      double *logabs, void *z1_out, void *stream) {
  return status3(vmcx::nnbf_forward(m, B, configs, workspace, ws_bytes, chunk, sign, logabs,
                                    z1_out, (hipStream_t)stream));
}

int vmcx_nnbf_forward_moved(const vmcx_nnbf_model *m, int64_t B, const int64_t *configs,
                            const int64_t *proposed, const void *z1, void *workspace,
                            int64_t ws_bytes, int64_t chunk, double *sign, double *logabs,
                            void *z1_proposed, void *stream) {
  return status3(vmcx::nnbf_forward_moved(m, B, configs, proposed, z1, workspace, ws_bytes,
                                          chunk, sign, logabs, z1_proposed,
                                          (hipStream_t)stream));
}

int vmcx_masked_rows(int64_t B, int64_t row_bytes, const uint8_t *mask, const void *src,
                     void *dst, void *stream) {
  return status3(vmcx::masked_rows(B, row_bytes, mask, src, dst, (hipStream_t)stream));
}

int vmcx_nnbf_sweep(const vmcx_nnbf_model *m, int64_t B, int n_edges,
                    const int32_t *edges_host, double hopping_rate, int n_sweeps,
                    uint64_t seed, int64_t sweep_offset, const double *u_choice,
                    const uint8_t *bits, const double *u_acc, int64_t *configs, void *z1,
                    double *sign, double *logabs, int64_t *n_accept, void *workspace,
                    int64_t ws_bytes, void *stream) {
  return status3(vmcx::nnbf_sweep(m, B, n_edges, edges_host, hopping_rate, n_sweeps, seed,
                                  sweep_offset, u_choice, bits, u_acc, configs, z1, sign,
                                  logabs, n_accept, workspace, ws_bytes, (hipStream_t)stream));
}

int vmcx_nnbf_local_energy(const vmcx_nnbf_model *m, int64_t B, int n_edges,
                           const int32_t *edges, double t, double U, const int64_t *configs,
                           const double *sign, const double *logabs, void *workspace,
                           int64_t ws_bytes, int64_t chunk, double *eloc, void *stream) {
  return status3(vmcx::nnbf_local_energy(m, B, n_edges, edges, t, U, configs, sign, logabs,
                                         workspace, ws_bytes, chunk, eloc,
                                         (hipStream_t)stream));
}

int vmcx_nnbf_logderiv(const vmcx_nnbf_model *m, int64_t B, const int64_t *configs,
                       void *workspace, int64_t ws_bytes, int64_t chunk, void *O, int o_dtype,
                       int64_t ld, double *sign, double *logabs, void *stream) {
  return status3(vmcx::nnbf_logderiv(m, B, configs, workspace, ws_bytes, chunk, O, o_dtype, ld,
                                     sign, logabs, (hipStream_t)stream));
}

int vmcx_nnbf_logderiv_factors(const vmcx_nnbf_model *m, int64_t B, const int64_t *configs,
                               void *workspace, int64_t ws_bytes, int64_t chunk, double *F,
                               int64_t ld, double *sign, double *logabs, void *stream) {
  return status3(vmcx::nnbf_logderiv_factors(m, B, configs, workspace, ws_bytes, chunk, F, ld,
                                             sign, logabs, (hipStream_t)stream));
}

int64_t vmcx_sr_factored_workspace_bytes(int64_t Ns, int n_factors, const int32_t *k,
                                         int n_blocks, const vmcx_sr_block *blocks) {
  return vmcx::sr_factored_workspace_bytes(Ns, n_factors, k, n_blocks, blocks);
}

int vmcx_sr_minsr_factored(int64_t Ns, int n_factors, const double *const *F,
                           const int64_t *ld, const int32_t *k, int n_blocks,
                           const vmcx_sr_block *blocks, int64_t Np, const double *E,
                           double energy_mean, double diag_shift, void *workspace,
                           int64_t ws_bytes, double *dp, double *stats, int32_t *info,
                           float *phase_ms, void *stream) {
  const int rc = vmcx::sr_minsr_factored(Ns, n_factors, F, ld, k, n_blocks, blocks, Np, E,
                                         energy_mean, diag_shift, workspace, ws_bytes, dp,
                                         stats, info, phase_ms, (hipStream_t)stream);
  return rc == 3 ? VMCX_ELIB : status_of(rc);
}

int vmcx_sr_factored_gradient(int64_t Ns, int n_factors, const double *const *F,
                              const int64_t *ld, const int32_t *k, int n_blocks,
                              const vmcx_sr_block *blocks, int64_t Np, const double *E,
                              double energy_mean, void *workspace, int64_t ws_bytes, double *g,
                              double *stats, void *stream) {
  const int rc = vmcx::sr_factored_gradient(Ns, n_factors, F, ld, k, n_blocks, blocks, Np, E,
                                            energy_mean, workspace, ws_bytes, g, stats,
                                            (hipStream_t)stream);
  return rc == 3 ? VMCX_ELIB : status_of(rc);
}

int vmcx_energy_stats(int64_t n, const double *E, double *out, void *stream) {
  return status_of(vmcx::energy_stats(n, E, out, (hipStream_t)stream));
}

int vmcx_to_bf16(int64_t n, const float *x, void *y, void *stream) {
  return status_of(vmcx::to_bf16(n, x, y, (hipStream_t)stream));
}

int vmcx_sgd_update(int64_t n, double *params, const double *dp, double lr,
                    void *stream) {
  return status_of(vmcx::sgd_update(n, params, dp, lr, (hipStream_t)stream));
}

int vmcx_lp_n_params(int C) { return vmcx::lp_n_params(C); }

int vmcx_lp_n_moves(int Lx, int Ly) { return vmcx::lp_n_moves(Lx, Ly); }

int vmcx_lp_logits(const float *params, int C, int Lx, int Ly, int64_t B,
                   const int64_t *configs, double *logits, int precision, void *stream) {
  return status_of(vmcx::lp_logits(params, C, Lx, Ly, B, configs, logits, precision,
                                   (hipStream_t)stream));
}

int vmcx_lp_propose(const float *params, int C, int Lx, int Ly, int64_t B,
                    const int64_t *configs, uint64_t seed, int64_t step, const double *u_choice,
                    double *logits_ws, int32_t *move, double *logq_fwd, int64_t *proposed,
                    double *logq_rev, int precision, void *stream) {
  return status_of(vmcx::lp_propose(params, C, Lx, Ly, B, configs, seed, step, u_choice,
                                    logits_ws, move, logq_fwd, proposed, logq_rev, precision,
                                    (hipStream_t)stream));
}

int vmcx_lp_accept(int64_t B, int n_sites, uint64_t seed, int64_t step, const double *u_acc,
                   const int32_t *move, const int64_t *proposed, const double *psign,
                   const double *plogabs, const double *logq_fwd, const double *logq_rev,
                   int64_t *configs, double *sign, double *logabs, int64_t *n_accept,
                   uint8_t *accepted, const int64_t *step_dev, void *stream) {
  return status_of(vmcx::lp_accept(B, n_sites, seed, step, u_acc, move, proposed, psign,
                                   plogabs, logq_fwd, logq_rev, configs, sign, logabs, n_accept,
                                   accepted, (hipStream_t)stream, step_dev));
}

int vmcx_counter_add(int64_t *counter, int64_t inc, void *stream) {
  return status_of(vmcx::counter_add(counter, inc, (hipStream_t)stream));
}

int64_t vmcx_lp_backward_workspace_bytes(int C, int64_t B) {
  return vmcx::lp_backward_workspace_bytes(C, B);
}

int vmcx_lp_backward(const float *params, int C, int Lx, int Ly, int64_t B,
                     const int64_t *configs, const double *dlogits, void *workspace,
                     int64_t ws_bytes, float *grad, void *stream) {
  return status_of(vmcx::lp_backward(params, C, Lx, Ly, B, configs, dlogits, workspace, ws_bytes,
                                     grad, (hipStream_t)stream));
}

int vmcx_lp_candidates(int64_t B, int Lx, int Ly, const int64_t *configs, int64_t *out,
                       void *stream) {
  return status_of(vmcx::lp_candidates(B, Lx, Ly, configs, out, (hipStream_t)stream));
}

int vmcx_lp_axpy(int64_t n, float a, const float *x, float *y, void *stream) {
  return status_of(vmcx::lp_axpy(n, a, x, y, (hipStream_t)stream));
}

int vmcx_lp_propose_given(const float *params, int C, int Lx, int Ly, int64_t B,
                          const int64_t *configs, const double *logits_x, uint64_t seed,
                          int64_t step, const double *u_choice, double *logits_p,
                          int32_t *move, double *logq_fwd, int64_t *proposed, double *logq_rev,
                          int precision, const int64_t *step_dev, void *stream) {
  return status_of(vmcx::lp_propose_given(params, C, Lx, Ly, B, configs, logits_x, seed, step,
                                          u_choice, logits_p, move, logq_fwd, proposed,
                                          logq_rev, precision, (hipStream_t)stream, step_dev));
}

int vmcx_lp_accept_commit(int64_t B, int n_sites, uint64_t seed, int64_t step,
                          const double *u_acc, const int32_t *move, const int64_t *proposed,
                          const double *psign, const double *plogabs, const double *logq_fwd,
                          const double *logq_rev, int64_t *configs, double *sign, double *logabs,
                          int64_t *n_accept, uint8_t *accepted, const int64_t *step_dev,
                          int n_moves, const double *logits_p, double *logits_x,
                          int64_t row_bytes, const void *z1p, void *z1, void *stream) {
  return status_of(vmcx::lp_accept_commit(B, n_sites, seed, step, u_acc, move, proposed, psign,
                                          plogabs, logq_fwd, logq_rev, configs, sign, logabs,
                                          n_accept, accepted, step_dev, n_moves, logits_p,
                                          logits_x, row_bytes, z1p, z1, (hipStream_t)stream));
}

int vmcx_lp_commit_logits(int64_t B, int n_moves, const uint8_t *accepted,
                          const double *logits_p, double *logits_x, void *stream) {
  return status_of(vmcx::lp_commit_logits(B, n_moves, accepted, logits_p, logits_x,
                                          (hipStream_t)stream));
}

int vmcx_sym_images(int64_t B, int n_sites, int n_group, const int64_t *configs,
                    const int32_t *perms, const int32_t *inv_perms, int64_t *images,
                    double *signs, void *stream) {
  return status_of(vmcx::sym_images(B, n_sites, n_group, configs, perms, inv_perms, images,
                                    signs, (hipStream_t)stream));
}

int64_t vmcx_peps_dense_size(const vmcx_peps_model *m) { return vmcx::peps_dense_size(m); }

int64_t vmcx_peps_workspace_bytes(const vmcx_peps_model *m, int64_t B, int64_t chunk,
                                  int n_edges) {
  return vmcx::peps_workspace_bytes(m, B, chunk, n_edges);
}

int64_t vmcx_peps_reuse_workspace_bytes(const vmcx_peps_model *m, int64_t B, int64_t chunk,
                                        int n_edges) {
  return vmcx::peps_reuse_workspace_bytes(m, B, chunk, n_edges);
}

int vmcx_peps_sweep_reuse(const vmcx_peps_model *m, int64_t B, int n_edges,
                          const int32_t *edges_host, double hopping_rate, int n_sweeps,
                          uint64_t seed, int64_t sweep_offset, const double *u_choice,
                          const uint8_t *bits, const double *u_acc, int64_t *configs,
                          double *sign, double *logabs, int64_t *n_accept, void *workspace,
                          int64_t ws_bytes, void *stream) {
  return status3(vmcx::peps_sweep_reuse(m, B, n_edges, edges_host, hopping_rate, n_sweeps, seed,
                                        sweep_offset, u_choice, bits, u_acc, configs, sign,
                                        logabs, n_accept, workspace, ws_bytes,
                                        (hipStream_t)stream));
}

int vmcx_peps_local_energy_reuse(const vmcx_peps_model *m, int64_t B, int n_edges,
                                 const int32_t *edges, double t, double U,
                                 const int64_t *configs, const double *sign,
                                 const double *logabs, void *workspace, int64_t ws_bytes,
                                 int64_t chunk, double *eloc, void *stream) {
  return status3(vmcx::peps_local_energy_reuse(m, B, n_edges, edges, t, U, configs, sign, logabs,
                                               workspace, ws_bytes, chunk, eloc,
                                               (hipStream_t)stream));
}

int vmcx_peps_forward(const vmcx_peps_model *m, int64_t B, const int64_t *configs,
                      void *workspace, int64_t ws_bytes, int64_t chunk, double *sign,
                      double *logabs, void *stream) {
  return status3(vmcx::peps_forward(m, B, configs, workspace, ws_bytes, chunk, sign, logabs,
                                    (hipStream_t)stream));
}

int vmcx_peps_sweep(const vmcx_peps_model *m, int64_t B, int n_edges, const int32_t *edges_host,
                    double hopping_rate, int n_sweeps, uint64_t seed, int64_t sweep_offset,
                    const double *u_choice, const uint8_t *bits, const double *u_acc,
                    int64_t *configs, double *sign, double *logabs, int64_t *n_accept,
                    void *workspace, int64_t ws_bytes, void *stream) {
  return status3(vmcx::peps_sweep(m, B, n_edges, edges_host, hopping_rate, n_sweeps, seed,
                                  sweep_offset, u_choice, bits, u_acc, configs, sign, logabs,
                                  n_accept, workspace, ws_bytes, (hipStream_t)stream));
}

int vmcx_peps_local_energy(const vmcx_peps_model *m, int64_t B, int n_edges,
                           const int32_t *edges, double t, double U, const int64_t *configs,
                           const double *sign, const double *logabs, void *workspace,
                           int64_t ws_bytes, int64_t chunk, double *eloc, void *stream) {
  return status3(vmcx::peps_local_energy(m, B, n_edges, edges, t, U, configs, sign, logabs,
                                         workspace, ws_bytes, chunk, eloc,
                                         (hipStream_t)stream));
}

int vmcx_peps_logderiv(const vmcx_peps_model *m, int64_t B, const int64_t *configs,
                       void *workspace, int64_t ws_bytes, int64_t chunk, void *O, int o_dtype,
                       int64_t ld, double *sign, double *logabs, void *stream) {
  return status3(vmcx::peps_logderiv(m, B, configs, workspace, ws_bytes, chunk, O, o_dtype, ld,
                                     sign, logabs, (hipStream_t)stream));
}

int vmcx_peps_logderiv_factors(const vmcx_peps_model *m, int64_t B, const int64_t *configs,
                               void *workspace, int64_t ws_bytes, int64_t chunk, double *F,
                               int64_t ld, double *sign, double *logabs, void *stream) {
  return status3(vmcx::peps_logderiv_factors(m, B, configs, workspace, ws_bytes, chunk, F, ld,
                                             sign, logabs, (hipStream_t)stream));
}

}  // extern "C"
