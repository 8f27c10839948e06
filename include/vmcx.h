/*
 * vmcx.h — C ABI of libvmcx.so, the MI355X (gfx950) kernels of the VMC inner
 * loop for the spinful Fermi-Hubbard model.
 *
 * Conventions
 *   - Every pointer argument is a DEVICE pointer owned by the caller (for
 *     example the data_ptr() of a torch tensor), unless stated otherwise.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Entry points only enqueue work; nothing synchronises the host.
 *   - Return value: VMCX_OK (0) or an error code; vmcx_last_error() returns
 *     the message of the last failure on the calling thread.
 *   - Configurations use the quimb site code {0 empty, 1 down, 2 up, 3 both},
 *     int64, row-major [B][n_sites]; site index x*Ly + y.
 *   - Orbitals use the netket order: up orbital of site i = i, down = N + i.
 *   - A Slater walker state is: configs, the inverse of the occupied-row
 *     matrix A = M[occ] (ainv, [B][NFP][NFP], column k of A^-1 stored
 *     contiguously at ainv + (b*NFP + k)*NFP), slot ([B][2N] uint8: the row
 *     of A that holds orbital i, 0xFF if i is empty), sign and ln|psi|
 *     ([B] float64) and the number of accepted moves since the last exact
 *     refresh ([B] int32).  NFP = vmcx_slater_nfp(n_f).
 *   - Limits: n_sites <= 64, n_f <= 64, 2*n_sites*NFP*8 bytes of M fit LDS.
 */
#ifndef VMCX_H
#define VMCX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VMCX_OK 0
#define VMCX_EINVAL 1
#define VMCX_EHIP 2
#define VMCX_ELIB 3

int vmcx_abi_version(void);
const char *vmcx_last_error(void);

/* Padded width (8, 16, 32 or 64) used for A^-1 storage; -1 if n_f > 64. */
int vmcx_slater_nfp(int n_f);

/*
 * Launch shapes picked by an autotuner (vmc_torch_amd/autotune.py): waves per
 * workgroup (1, 2, 4 or 8; 0 = the built-in default 4) of a kernel family
 * for a key.  kind 0 = the Slater sweep, 1 = the Slater local energy, 2 =
 * the Slater refresh; key = 2 NFP + (1 for the FP32 form).  Process-wide,
 * read at every launch.  No counterpart in the reference.
 */
int vmcx_set_launch_waves(int kind, int key, int waves);
int vmcx_get_launch_waves(int kind, int key);

/*
 * Exact (re)build of the Slater walker state from configs and M, by an
 * in-register Gauss-Jordan inversion with partial pivoting (FP64).
 *   ln|psi| and sign of psi(x) = det(M[occ ascending, :]).
 *   ainv / slot / since_refresh may be NULL: then only sign and logabs are
 *   written (batched log-amplitude of arbitrary configurations).
 *   only_if_since >= 0: rebuild only walkers with since_refresh >= it.
 * Replaces: SlaterDeterminant_GPU.amplitude + _base.log_amplitude
 *   (GPU/models/slater.py:36-61, GPU/models/_base.py:101-109) and the
 *   forward_log calls of GPU/sampler.py:239 and GPU/vmc_utils.py:407-432.
 */
int vmcx_slater_refresh(int64_t B, int n_sites, int n_f, const double *M,
                        const int64_t *configs, double *ainv, uint8_t *slot,
                        double *sign, double *logabs, int32_t *since_refresh,
                        int only_if_since, void *stream);

/*
 * n_sweeps Metropolis sweeps over the edge list (exchange / hop proposals,
 * Sherman-Morrison rank-1 / rank-2 ratio and inverse updates).
 *   edges: [n_edges][2] int32, sweep order (row edges then column edges).
 *   Draws: if u_choice is NULL, Philox4x32-10 with key = seed and counter
 *   (edge, sweep_offset + sweep, walker, 0); else injected arrays of shape
 *   [n_sweeps][n_edges][B]: u_choice, u_acc (float64), bits (uint8,
 *   bit0 = the (0,3)/(3,0) choice, bit1 = the (1,2)/(2,1) choice).
 *   n_accept ([B] int64, may be NULL) is incremented per accepted move.
 *   refresh_every > 0: after each sweep, walkers with since_refresh >=
 *   refresh_every are rebuilt exactly (vmcx_slater_refresh semantics).
 *   sweep_offset_dev (may be NULL): a device int64 added to sweep_offset when
 *   the kernel runs, so a captured graph advances its draws between replays
 *   (with vmcx_counter_add).
 * Replaces: MetropolisExchangeSpinfulSamplerGPU.step / burn_in
 *   (GPU/sampler.py:175-197, 216-388) with propose_exchange_or_hopping_vec
 *   (GPU/sampler.py:47-137).
 */
int vmcx_slater_sweep(int64_t B, int n_sites, int n_f, const double *M,
                      int n_edges, const int32_t *edges, double hopping_rate,
                      int n_sweeps, uint64_t seed, int64_t sweep_offset,
                      const double *u_choice, const uint8_t *bits,
                      const double *u_acc, int refresh_every,
                      int64_t *configs, double *ainv, uint8_t *slot,
                      double *sign, double *logabs, int32_t *since_refresh,
                      int64_t *n_accept, const int64_t *sweep_offset_dev, void *stream);

/* counter[0] += inc on `stream` (a one-thread kernel: graph-capturable, no
 * memset node).  Advances the device draw counters of vmcx_slater_sweep and
 * vmcx_lp_propose_given / vmcx_lp_accept inside a captured graph.  No
 * counterpart in the reference (its draws come from torch's host-seeded
 * generator). */
int vmcx_counter_add(int64_t *counter, int64_t inc, void *stream);

/*
 * Hubbard local energy of every walker from its Slater state:
 *   E_loc = U * n_double + sum_{hops} (-t) * s_JW * psi(eta)/psi(x),
 *   s_JW = the symmray-order Jordan-Wigner sign times the netket row
 *   reordering sign, both popcount parities; psi(eta)/psi(x) from A^-1.
 *   edges: [n_edges][2] int32 (any order).
 * Replaces: _evaluate_energy_impl (GPU/vmc_utils.py:322-476) with
 *   get_conn_batch_gpu (hamiltonian_torch.py:586-655).
 */
int vmcx_slater_local_energy(int64_t B, int n_sites, int n_f, const double *M,
                             int n_edges, const int32_t *edges, double t,
                             double U, const int64_t *configs,
                             const double *ainv, const uint8_t *slot,
                             double *eloc, void *stream);

/*
 * Per-sample log-derivatives O[b][i*n_f + j] = d ln|psi| / d M[i][j]
 * = A^-1[j][slot(i)] for occupied i, else 0.  o_dtype: 0 = float64,
 * 1 = float32.  ld = row stride of O in elements (>= 2*n_sites*n_f).
 * Replaces: compute_grads_gpu with use_log_amp=True for the Slater model
 *   (GPU/vmc_utils.py:1389-1634).
 */
int vmcx_slater_logderiv(int64_t B, int n_sites, int n_f, const double *ainv,
                         const uint8_t *slot, void *O, int o_dtype, int64_t ld,
                         void *stream);

/*
 * The FP32 form of the four Slater entry points above (the precision form of
 * the 8x8 Slater configuration).  The walker's A^-1 is stored in FP32
 * ([B][NFP][NFP] float, same layout); the sweep forms ratios, determinant
 * ratios and Sherman-Morrison / Woodbury updates in FP32 from M32 (M rounded
 * to FP32, the same [2N][n_f] layout), compares the acceptance uniform with
 * the squared ratio and accumulates sign and ln|psi| in FP64; the exact
 * refresh (vmcx_slater_refresh_f32, and the refresh between sweeps) runs
 * the FP64 Gauss-Jordan on the FP64 M and rounds A^-1 to FP32.  The local
 * energy forms each ratio in FP32 and sums in FP64.  Draws, proposal order,
 * acceptance rule and outputs otherwise as the FP64 entry points.
 * Replaces: SlaterDeterminant_GPU with dtype=torch.float32
 *   (GPU/models/slater.py:29-34) on the path of the FP64 entry points.
 */
int vmcx_slater_refresh_f32(int64_t B, int n_sites, int n_f, const double *M,
                            const int64_t *configs, float *ainv, uint8_t *slot, double *sign,
                            double *logabs, int32_t *since_refresh, int only_if_since,
                            void *stream);
int vmcx_slater_sweep_f32(int64_t B, int n_sites, int n_f, const double *M, const float *M32,
                          int n_edges, const int32_t *edges, double hopping_rate, int n_sweeps,
                          uint64_t seed, int64_t sweep_offset, const double *u_choice,
                          const uint8_t *bits, const double *u_acc, int refresh_every,
                          int64_t *configs, float *ainv, uint8_t *slot, double *sign,
                          double *logabs, int32_t *since_refresh, int64_t *n_accept,
                          const int64_t *sweep_offset_dev, void *stream);
int vmcx_slater_local_energy_f32(int64_t B, int n_sites, int n_f, const float *M32,
                                 int n_edges, const int32_t *edges, double t, double U,
                                 const int64_t *configs, const float *ainv,
                                 const uint8_t *slot, double *eloc, void *stream);
int vmcx_slater_logderiv_f32(int64_t B, int n_sites, int n_f, const float *ainv,
                             const uint8_t *slot, void *O, int o_dtype, int64_t ld,
                             void *stream);
/* params[i] -= lr * dp[i] for FP32 parameters (the product and difference in
 * FP64, rounded once).  Replaces: SGDGPU.step (GPU/optimizer.py:144-186) on
 * an FP32 model. */
int vmcx_sgd_update_f32(int64_t n, float *params, const double *dp, double lr, void *stream);

/*
 * Batched connected configurations of the Hubbard Hamiltonian in the
 * reference's term-major order: terms = for each edge (+1 spin, -1 spin)
 * then one on-site term per site.
 *   Pass 1 (vmcx_hubbard_conn_count): counts[n_terms] int64.
 *   Pass 2 (vmcx_hubbard_conn_fill): offsets[n_terms] int64 = exclusive
 *   scan of counts; writes etas [T][n_sites] int64, coeffs [T] float64,
 *   batch_ids [T] int64.
 * Replaces: SpinfulFermionGPUMixin.get_conn_batch_gpu
 *   (hamiltonian_torch.py:586-655).
 */
int vmcx_hubbard_conn_count(int64_t B, int n_sites, int n_edges,
                            const int32_t *edges, const int64_t *configs,
                            int64_t *counts, void *stream);
int vmcx_hubbard_conn_fill(int64_t B, int n_sites, int n_edges,
                           const int32_t *edges, double t, double U,
                           const int64_t *configs, const int64_t *offsets,
                           int64_t *etas, double *coeffs, int64_t *batch_ids,
                           void *stream);

/*
 * Stochastic reconfiguration in sample space (minSR):
 *   Ot = (O - mean_s O)/sqrt(Ns), eps = (E - E_mean)/sqrt(Ns),
 *   T = Ot Ot^T + diag_shift I, alpha = T^-1 eps, dp = Ot^T alpha.
 *   O: [Ns][ldo] row-major per-sample log-derivatives, float64 (o_dtype 0)
 *   or float32 (o_dtype 1).  flags bit 0: centre O in place (O is
 *   overwritten by Ot, saving an Ns x Np copy).  E: [Ns] float64.
 *   energy_mean: the mean used to centre E; NaN = the sample mean.
 *   gram_dtype: 0 = FP64 Gram (SYRK; needs FP64 O); 1 = FP32 and 2 = BF16
 *   operands with FP32 accumulation; 3 = three BF16 products of the split
 *   Ot = hi + lo (hi hi^T + hi lo^T + lo hi^T, ~2^-16 operand rounding).
 *   The shifted Gram is factorised by a blocked FP64 Cholesky.  max_iter > 0:
 *   preconditioned conjugate gradients on the exact operator Ot Ot^T + shift
 *   (FP64 accumulation), with that factor as the preconditioner, until
 *   ||res|| <= tol ||eps||; a reduced-precision Gram that is not positive
 *   definite at the shift gets its preconditioner shift raised tenfold (up to
 *   six times) while the operator keeps diag_shift.
 *   workspace: >= vmcx_sr_workspace_bytes(Ns, Np, o_dtype, gram_dtype, flags).
 *   stats (may be NULL): [5] float64 = mean E, var E, ||T alpha - eps|| /
 *   ||eps||, ||g|| (g = Ot^T eps, the energy gradient), PCG iterations.
 *   info (may be NULL): Cholesky info of the factor used (0 = success, k > 0
 *   = leading minor k not positive).
 *   phase_ms (HOST pointer, may be NULL): [6] float = milliseconds of the
 *   centring, Gram, factorisation, solve + PCG and dp / statistics phases,
 *   and [5] the mean duration of one diagonal-block launch of the
 *   factorisation, all timed with HIP events on `stream`; when given, the
 *   call waits for the stream before returning.
 * Replaces: minSR_solver_gpu (GPU/optimizer.py:471-614; MinSRGPU.solve
 *   :873-895), whose LU solve becomes a Cholesky solve of the same SPD system.
 */
int64_t vmcx_sr_workspace_bytes(int64_t Ns, int64_t Np, int o_dtype, int gram_dtype,
                                int flags);
int vmcx_sr_minsr(int64_t Ns, int64_t Np, void *O, int o_dtype, int64_t ldo, int flags,
                  const double *E, double energy_mean, double diag_shift, int gram_dtype,
                  int max_iter, double tol, void *workspace, int64_t ws_bytes, double *dp,
                  double *stats, int32_t *info, float *phase_ms, void *stream);

/*
 * minSR on a factored O: the same step as vmcx_sr_minsr with gram_dtype 0
 * (FP64 Gram, Cholesky solve), for models whose per-sample log-derivative is
 * made of blocks that are a factor or an outer product of two factors (the
 * dense layers of an MLP: dL/dW = delta x^T).  Factor f is the row-major
 * FP64 matrix F[f] [Ns][ld[f]] of width k[f] (F: HOST array of device
 * pointers; n_factors <= 8).  Block b (n_blocks <= 8):
 *   b_factor < 0:  O[s][w_offset + i]         = scale F_a[s][i]
 *   b_factor >= 0: O[s][w_offset + i k_b + j] = scale F_a[s][i] F_b[s][j]
 *                  and, bias_offset >= 0, O[s][bias_offset + i] = scale F_a[s][i]
 * (a = a_factor, b = b_factor).  The blocks must tile [0, Np) exactly.
 * Then O O^T = sum over blocks of scale^2 G_a o (G_b [+ 1]) with G_f = F_f F_f^T
 * (o: elementwise product; + 1 for a bias block): the factor Grams are formed
 * on the FP64 matrix cores with inner dimension k[f] instead of Np.  The
 * centred Gram is P (O O^T) P / Ns (P = I - 1 1^T / Ns, double centring),
 * dp = O^T P alpha / sqrt(Ns), one FP64 GEMM per outer-product block.  No
 * Ns x Np matrix is formed.
 *   stats / info as vmcx_sr_minsr (iterations 0; the residual is that of the
 *   Cholesky solve against the FP64 Gram).  phase_ms (HOST, may be NULL): [6]
 *   float = factor Grams, combination + centring, factorisation, solve,
 *   dp / statistics, [5] mean diagonal-block launch of the factorisation.
 * Replaces: minSR_solver_gpu (GPU/optimizer.py:471-614) for models built
 *   from dense layers (NNBF.py:36-162), with the Gram of GPU/optimizer.py:530-536
 *   formed from the factors.
 */
typedef struct {
  int32_t a_factor, b_factor;
  double scale;
  int64_t w_offset, bias_offset;
} vmcx_sr_block;
int64_t vmcx_sr_factored_workspace_bytes(int64_t Ns, int n_factors, const int32_t *k,
                                         int n_blocks, const vmcx_sr_block *blocks);
int vmcx_sr_minsr_factored(int64_t Ns, int n_factors, const double *const *F,
                           const int64_t *ld, const int32_t *k, int n_blocks,
                           const vmcx_sr_block *blocks, int64_t Np, const double *E,
                           double energy_mean, double diag_shift, void *workspace,
                           int64_t ws_bytes, double *dp, double *stats, int32_t *info,
                           float *phase_ms, void *stream);

/*
 * The energy gradient on a factored O (the blocks of vmcx_sr_minsr_factored):
 *   g = (1/Ns) sum_s (E_s - energy_mean) O_s  (energy_mean NaN = sample mean),
 * without the Gram or the factorisation.  workspace: the same size as
 * vmcx_sr_minsr_factored's.  stats (may be NULL): [5] = mean E, var E, 0,
 * ||g||, 0.
 * Replaces: the run_SR=False branch of minSR_solver_gpu
 *   (GPU/optimizer.py:581-614) for models built from dense layers.
 */
int vmcx_sr_factored_gradient(int64_t Ns, int n_factors, const double *const *F,
                              const int64_t *ld, const int32_t *k, int n_blocks,
                              const vmcx_sr_block *blocks, int64_t Np, const double *E,
                              double energy_mean, void *workspace, int64_t ws_bytes, double *g,
                              double *stats, void *stream);

/*
 * out[0] = mean of E[0..n), out[1] = its variance (1/n normalisation), both
 * from compensated (Neumaier) sums, the variance as the mean of (E - mean)^2.
 * Replaces: VMC_GPU.compute_global_energy_stats (GPU/VMC.py:513-534) on one
 *   device, whose sums are naive.
 */
int vmcx_energy_stats(int64_t n, const double *E, double *out, void *stream);

/*
 * FP64 symmetric rank-K update on the matrix cores (a building block of the
 * minSR Gram and of the blocked Cholesky, exported for direct use):
 *   C = beta C + alpha M M^T on the lower triangle of the n x n column-major
 *   C (ldc >= n); the strictly upper triangle is not touched.
 *   M is n x K: rowmajor = 1: M(i, k) = X[i*ldx + k] (ldx >= K);
 *   rowmajor = 0: M(i, k) = X[k*ldx + i] (ldx >= n).
 *   splits > 1: the reduction is cut into `splits` equal chunks; chunk s
 *   writes beta C_s + alpha M_s M_s^T into C + s*cstride (cstride >= ldc*n).
 * Replaces: torch.matmul(O, O.T) of the minSR Gram (GPU/optimizer.py:530-536)
 *   and the trailing update of the LU/Cholesky that follows it.
 */
int vmcx_syrk_f64(int64_t n, int64_t K, const double *X, int64_t ldx, int rowmajor,
                  double alpha, double beta, double *C, int64_t ldc, int splits,
                  int64_t cstride, void *stream);

/*
 * Stochastic reconfiguration in parameter space by MINRES:
 *   (S + diag_shift I) dp = g,  S x = O^T (O x) / Ns - (Obar . x) Obar,
 *   g = (1/Ns) sum_s (E_s - energy_mean) O_s,  Obar = mean_s O_s,
 * S is never formed (two passes over O per iteration).  Paige-Saunders
 * MINRES with the convergence test |r| < rtol ||A|| ||dp|| (running
 * estimates, as scipy.sparse.linalg.minres); at most maxiter iterations.
 * The whole loop runs on `stream` without host synchronisation.
 *   O: [Ns][ldo] float64 (o_dtype 0) or float32 (o_dtype 1); E: [Ns].
 *   energy_mean: NaN = the sample mean.
 *   workspace: >= vmcx_sr_minres_workspace_bytes(Ns, Np).
 *   dp: [Np] float64 out; maxiter = 0 returns dp = g (the reference's
 *   run_SR=False).  stats (may be NULL): [5] = mean E, var E, residual
 *   estimate |phibar|, ||g||, iterations.  info (may be NULL): 0 =
 *   converged (or g = 0), maxiter = not converged.
 * Replaces: distributed_minres_solver_gpu + torch_minres
 *   (GPU/optimizer.py:257-465, GPU/torch_utils.py:747-857;
 *   DistributedSRMinresGPU.solve :846-882) on one device.
 */
int64_t vmcx_sr_minres_workspace_bytes(int64_t Ns, int64_t Np);
int vmcx_sr_minres(int64_t Ns, int64_t Np, const void *O, int o_dtype, int64_t ldo,
                   const double *E, double energy_mean, double diag_shift, double rtol,
                   int maxiter, void *workspace, int64_t ws_bytes, double *dp, double *stats,
                   int32_t *info, void *stream);

/*
 * vmcx_sr_minres without a workspace argument and without any allocation:
 * no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The workspace is a fixed 20 MiB device array inside the
 * library, one per GPU, held until the process ends:
 * vmcx_sr_minres_workspace_bytes(Ns, Np) must not exceed
 * vmcx_sr_minres_noalloc_max_bytes() (e.g. Np <= 16384 with Ns <= 65536),
 * otherwise VMCX_EINVAL.  Arguments and results as vmcx_sr_minres.  Calls
 * that overlap on one GPU (two streams, two threads) share that array and
 * race; calls in sequence on one stream are safe (no entry is read before the
 * same call writes it).
 */
int64_t vmcx_sr_minres_noalloc_max_bytes(void);
int vmcx_sr_minres_noalloc(int64_t Ns, int64_t Np, const void *O, int o_dtype, int64_t ldo,
                           const double *E, double energy_mean, double diag_shift, double rtol,
                           int maxiter, double *dp, double *stats, int32_t *info,
                           void *stream);

/*
 * Neural-network backflow (NNBF) model: two hidden layers of width `hidden`.
 *   psi(x) = det((M + bf_scale * reshape(Wout act(W2 act(W1 n + b1) + b2)
 *                 + bout, (2N, n_f)))[occupied rows ascending])
 * Weights in the reference's layouts (row-major [out][in]).  gemm = 0: FP32
 * device copies, FP32 GEMMs; gemm = 1: the two dense layers on BF16 copies of
 * W2 and Wout (and BF16 activations) with FP32 accumulation; gemm = 2: the
 * FP64 weights (the *_64 pointers) and FP64 arithmetic throughout, the
 * reference's float64 model.  The walker state z1 ([B][hidden]) is float for
 * gemm 0 / 1 and double for gemm 2.  In BF16 mode (path 0) the forward of
 * configurations (vmcx_nnbf_forward, the connected configurations of
 * vmcx_nnbf_local_energy) runs both dense layers as the library's own
 * matrix-core GEMMs and forms only the occupied orbital rows of the output
 * layer (the determinant's matrix, FP32) instead of the whole backflow
 * correction; path 1 keeps the rocBLAS GEMMs (sweeps and log-derivatives use
 * them in either path).
 * Replaces: NNBF_GPU (GPU/models/NNBF.py:36-162).
 */
typedef struct vmcx_nnbf_model {
  int32_t n_sites;  /* N <= 64 */
  int32_t n_f;      /* electrons <= 64 */
  int32_t hidden;   /* width of both hidden layers */
  int32_t act;      /* 0 = ReLU, 1 = tanh */
  int32_t gemm;     /* 0 = FP32 GEMMs, 1 = BF16 operands, FP32 accumulation, 2 = FP64 */
  int32_t path;     /* gemm 1: 0 = the library's own matrix-core layers, 1 = rocBLAS */
  double bf_scale;
  const float *M, *W1, *b1, *W2, *b2, *Wout, *bout;   /* gemm 0 / 1 */
  const void *W2_bf16, *Wout_bf16;                     /* gemm 1 */
  const double *M64, *W1_64, *b1_64, *W2_64, *b2_64, *Wout_64, *bout_64;  /* gemm 2 */
  int32_t det_fp32;   /* gemm 1, path 0: determinants of the amplitude-only passes
                         (forward, local energy, proposals) eliminated in FP32 (the
                         mixed-precision form); A^-1 (sweeps, O_k) stays FP64 */
  int32_t reserved2;
} vmcx_nnbf_model;

/* Workspace bytes for B walkers, forward chunks of `chunk` configurations
 * and n_edges lattice edges (local energy); sweeps need chunk >= B. */
int64_t vmcx_nnbf_workspace_bytes(const vmcx_nnbf_model *m, int64_t B, int64_t chunk,
                                  int n_edges);

/* Batched (sign, ln|psi|) of B configurations; z1_out (may be NULL) receives
 * the first-layer pre-activations [B][hidden] (the sweep's walker state).
 * Replaces: NNBF_GPU.forward_log via _base.vamp_log (GPU/models/_base.py:111-130). */
int vmcx_nnbf_forward(const vmcx_nnbf_model *m, int64_t B, const int64_t *configs,
                      void *workspace, int64_t ws_bytes, int64_t chunk, double *sign,
                      double *logabs, void *z1_out, void *stream);

/* (sign, ln|psi|) of B proposed configurations `proposed`, row b differing
 * from configs[b] by a few moved electrons, with the first layer formed from
 * z1 (configs' pre-activations [B][hidden], as vmcx_nnbf_forward's z1_out)
 * plus, for each orbital whose occupation differs, +-W1[:, orbital] (+ when
 * occupied in `proposed`); the rest of the forward as vmcx_nnbf_forward.
 * z1_proposed receives the proposed configurations' pre-activations.
 * Replaces: the model amplitude of the proposed configurations in a
 * sampler step (GPU/sampler.py:216-388 with an NNBF model) in incremental
 * form. */
int vmcx_nnbf_forward_moved(const vmcx_nnbf_model *m, int64_t B, const int64_t *configs,
                            const int64_t *proposed, const void *z1, void *workspace,
                            int64_t ws_bytes, int64_t chunk, double *sign, double *logabs,
                            void *z1_proposed, void *stream);

/* dst rows (row_bytes each, a multiple of 4) replaced by src rows where
 * mask[b] != 0: the walker state of the accepted proposals.
 * Replaces: the accepted-row updates of a sampler step
 * (GPU/sampler.py:353-366, `x[accept_mask] = proposed[accept_mask]`). */
int vmcx_masked_rows(int64_t B, int64_t row_bytes, const uint8_t *mask, const void *src,
                     void *dst, void *stream);

/* Metropolis sweeps for NNBF walkers: per edge, all walkers propose (same
 * proposal table and draws as vmcx_slater_sweep), the dense layers and the
 * determinant are evaluated for the proposed configurations, and accepted
 * moves update configs, z1 ([B][hidden], float or double as above), sign and
 * logabs in place.
 * edges_host: HOST array [n_edges][2] int32.
 * Replaces: MetropolisExchangeSpinfulSamplerGPU.step with an NNBF model
 * (GPU/sampler.py:216-388). */
int vmcx_nnbf_sweep(const vmcx_nnbf_model *m, int64_t B, int n_edges,
                    const int32_t *edges_host, double hopping_rate, int n_sweeps,
                    uint64_t seed, int64_t sweep_offset, const double *u_choice,
                    const uint8_t *bits, const double *u_acc, int64_t *configs, void *z1,
                    double *sign, double *logabs, int64_t *n_accept, void *workspace,
                    int64_t ws_bytes, void *stream);

/* Local energies: walker-major connected configurations, their NNBF
 * amplitudes in chunks, and E_loc = U n_double + sum coeff psi'/psi.
 * Synchronises the stream once (to size the connected set).
 * Replaces: _evaluate_energy_impl (GPU/vmc_utils.py:322-476) for NNBF. */
int vmcx_nnbf_local_energy(const vmcx_nnbf_model *m, int64_t B, int n_edges,
                           const int32_t *edges, double t, double U, const int64_t *configs,
                           const double *sign, const double *logabs, void *workspace,
                           int64_t ws_bytes, int64_t chunk, double *eloc, void *stream);

/* Per-sample d ln|psi| / d theta, parameter order [M, W1, b1, W2, b2, Wout,
 * bout] (NNBF.py:44-48); O row stride ld; o_dtype 0 = float64, 1 = float32.
 * Also writes sign / logabs.  Replaces: compute_grads_gpu for NNBF
 * (GPU/vmc_utils.py:1389-1634). */
int vmcx_nnbf_logderiv(const vmcx_nnbf_model *m, int64_t B, const int64_t *configs,
                       void *workspace, int64_t ws_bytes, int64_t chunk, void *O, int o_dtype,
                       int64_t ld, double *sign, double *logabs, void *stream);

/* The factors of the per-sample log-derivative instead of O: row s of F
 * ([B][ld] float64, ld >= 2N nf + 2N + 4 hidden) is [gM | n | dz1 | h1 |
 * dz2 | h2]: gM (2N x nf) = d ln|psi| / dM, n the occupations, dz1 / dz2 the
 * back-propagated errors of the hidden layers, h1 / h2 their activations.
 * O = [gM, dz1 n^T, dz1, dz2 h1^T, dz2, s gM h2^T, s gM] (s = bf_scale), the
 * blocks vmcx_sr_minsr_factored consumes.  Also writes sign / logabs.
 * Replaces: compute_grads_gpu for NNBF (GPU/vmc_utils.py:1389-1634) when the
 * SR step takes the factored form. */
int vmcx_nnbf_logderiv_factors(const vmcx_nnbf_model *m, int64_t B, const int64_t *configs,
                               void *workspace, int64_t ws_bytes, int64_t chunk, double *F,
                               int64_t ld, double *sign, double *logabs, void *stream);

/* y (bf16) = x (fp32), n elements (weight copies for gemm = 1). */
int vmcx_to_bf16(int64_t n, const float *x, void *y, void *stream);

/* params[i] -= lr * dp[i].  Replaces: SGDGPU.step (GPU/optimizer.py:144-186). */
int vmcx_sgd_update(int64_t n, double *params, const double *dp, double lr,
                    void *stream);

/*
 * Z2-symmetric fermionic PEPS (fPEPS), optionally with the NN-fTNS backflow
 * theta(x) = theta + eta * MLP(n(x)) (hidden > 0).
 *   Site tensors T_s[p][l][u][d][r] (p = 0, 1 down, 2 up, 3 both; l/u/d/r of
 *   dimension D, 1 on the open boundary; virtual index k has parity
 *   k >= D/2), Z2-even: only entries with even total parity are parameters.
 *   The flat parameter vector theta lists the allowed entries site by site
 *   (site i = x*Ly + y) in row-major (p, l, u, d, r) order; dense_map
 *   [sum_s 4 dl du dd dr] (device, int32) maps each dense position, in the
 *   same order, to its theta index or -1.
 *   Amplitude, in the Hamiltonian's Jordan-Wigner order (sites row-major,
 *   down before up):  psi(x) = (-1)^{C(P,2)} contract_s S_s with
 *   S_s[l,u,d,r] = T_s[x_s,l,u,d,r] (-1)^{q(d) (q(r) + P_right(s))}
 *   (P: singly occupied sites; P_right(s): those right of s in its row), the
 *   Grassmann tensor network with leg order (l, u, d, r).  Contracted
 *   exactly (2 <= Lx <= 4, Ly <= 8, D <= 4: boundary bonds up to D^2).
 *   MLP (hidden > 0): n = [up | down] occupations (2N), two hidden layers
 *   act(W1 n + b1), act(W2 h + b2), output y = Wo h + bo (n_params), FP64,
 *   row-major [out][in] weights; theta(x) = theta + eta y (mode 0, the
 *   reference's additive backflow) or theta * (1 + eta y) (mode 1, the
 *   multiplicative site-tensor correction; no reference form).
 *   O layout (vmcx_peps_logderiv): [theta (n_params) | W1 | b1 | W2 | b2 | Wo
 *   | bo] (the MLP part only with hidden > 0).
 * Replaces: fPEPS_Model_GPU (GPU/models/pureTNS.py:24-156) and
 *   NNfTNS_Model_GPU._apply_backflow / _contract_tn (GPU/models/NNfTNS.py:
 *   280-955) for exact contraction; the quimb/symmray block layout is
 *   replaced by the dense-map layout above.
 */
typedef struct vmcx_peps_model {
  int32_t Lx, Ly, D, n_params;
  const int32_t *dense_map;
  const double *theta;
  int32_t hidden; /* 0 = plain fPEPS */
  int32_t act;    /* 0 = ReLU, 1 = tanh */
  double eta;
  const double *W1, *b1, *W2, *b2, *Wo, *bo;
  int32_t mode;   /* 0 = additive theta + eta y, 1 = multiplicative theta (1 + eta y) */
  int32_t amp_fp32; /* 1: amplitudes (forward, sweep proposals, connected configurations of
                       the local energy) contracted in FP32 on the matrix cores; the
                       log-derivative and the environment-reuse entry points stay FP64 */
  int32_t chi;      /* max_bond of fPEPS_Model_GPU (pureTNS.py:81-146).  <= 0: exact
                       contraction (Lx <= 4, Ly <= 8).  > 0 and below D^rows of a side (or a
                       lattice beyond the exact kernels): chi-truncated boundary MPS, rows
                       0 .. Lx/2-1 from the top and Lx-1 .. Lx/2 from the bottom, each
                       truncated left to right onto the chi leading Schmidt vectors of every
                       cut after each absorbed row; chi <= 32, max(chi, D) * D <= 64,
                       Lx * Ly <= 64, Ly <= 16; FP64 amplitudes only (forward, sweep, local
                       energy) */
  int32_t bmps_nocache; /* with chi-truncation (plain fPEPS): 0 = sweeps keep every walker's
                           boundary after each row and recompute a proposal's boundaries
                           from its first changed row on (bit-identical to the full
                           contraction); 1 = recompute every proposal from scratch */
} vmcx_peps_model;

/* Number of dense tensor positions (length of dense_map); -1 on bad sizes. */
int64_t vmcx_peps_dense_size(const vmcx_peps_model *m);
/* Workspace bytes for B walkers, chunks of `chunk` configurations and
 * n_edges edges (local energy); sweeps need chunk >= B. */
int64_t vmcx_peps_workspace_bytes(const vmcx_peps_model *m, int64_t B, int64_t chunk,
                                  int n_edges);
/* (sign, ln|psi|) of B configurations.
 * Replaces: fPEPS_Model_GPU.log_amplitude / vamp_log (pureTNS.py:109-142). */
int vmcx_peps_forward(const vmcx_peps_model *m, int64_t B, const int64_t *configs,
                      void *workspace, int64_t ws_bytes, int64_t chunk, double *sign,
                      double *logabs, void *stream);
/* Metropolis sweeps (proposal table, draws and acceptance as vmcx_slater_sweep),
 * full amplitude per proposal.  edges_host: HOST [n_edges][2] int32.
 * Replaces: MetropolisExchangeSpinfulSamplerGPU.step with an fPEPS / NN-fTNS
 * model (GPU/sampler.py:216-388). */
int vmcx_peps_sweep(const vmcx_peps_model *m, int64_t B, int n_edges, const int32_t *edges_host,
                    double hopping_rate, int n_sweeps, uint64_t seed, int64_t sweep_offset,
                    const double *u_choice, const uint8_t *bits, const double *u_acc,
                    int64_t *configs, double *sign, double *logabs, int64_t *n_accept,
                    void *workspace, int64_t ws_bytes, void *stream);
/* E_loc of B walkers (walker-major connected configurations, amplitudes in
 * chunks).  Replaces: _evaluate_energy_impl (GPU/vmc_utils.py:322-476). */
int vmcx_peps_local_energy(const vmcx_peps_model *m, int64_t B, int n_edges,
                           const int32_t *edges, double t, double U, const int64_t *configs,
                           const double *sign, const double *logabs, void *workspace,
                           int64_t ws_bytes, int64_t chunk, double *eloc, void *stream);
/*
 * Environment reuse (plain fPEPS, hidden = 0): each walker keeps its column
 * environments L_y (columns < y contracted) and R_y (columns >= y), and a
 * configuration that differs from the walker only in columns c0 .. c1 is
 * contracted from L_c0 and R_{c1+1} through those columns alone.  The local
 * signs of the sites left of c0 change with the hop (S_s carries
 * (-1)^{q(d) P_right(s)}); on L_c0 that change is a sign times a parity on
 * the bond indices of the affected rows, applied in the kernel.  Results
 * equal vmcx_peps_sweep / vmcx_peps_local_energy up to rounding (same
 * proposals, draws and tests).  NN-fTNS models are refused (theta(x) changes
 * with every configuration).
 */
/* Workspace bytes of the reuse entry points (sweeps: chunk >= B, n_edges 0). */
int64_t vmcx_peps_reuse_workspace_bytes(const vmcx_peps_model *m, int64_t B, int64_t chunk,
                                        int n_edges);
/* Metropolis sweeps as vmcx_peps_sweep, proposals from the cached
 * environments; stale environments (columns changed by accepted moves) are
 * rebuilt incrementally when an edge needs them.
 * Replaces: MetropolisExchangeSpinfulSamplerReuse_GPU.step (GPU/sampler.py:
 * 391-700) with fPEPS_Model_reuse_GPU.forward_reuse_log (GPU/models/
 * pureTNS.py:1366-1455) and the environment caches / updates (:1456-2228). */
int vmcx_peps_sweep_reuse(const vmcx_peps_model *m, int64_t B, int n_edges,
                          const int32_t *edges_host, double hopping_rate, int n_sweeps,
                          uint64_t seed, int64_t sweep_offset, const double *u_choice,
                          const uint8_t *bits, const double *u_acc, int64_t *configs,
                          double *sign, double *logabs, int64_t *n_accept, void *workspace,
                          int64_t ws_bytes, void *stream);
/* E_loc as vmcx_peps_local_energy: the walkers' environments once, each
 * connected configuration through the one or two columns its hop touches.
 * Replaces: evaluate_energy_reuse (GPU/vmc_utils.py:584-1320). */
int vmcx_peps_local_energy_reuse(const vmcx_peps_model *m, int64_t B, int n_edges,
                                 const int32_t *edges, double t, double U,
                                 const int64_t *configs, const double *sign,
                                 const double *logabs, void *workspace, int64_t ws_bytes,
                                 int64_t chunk, double *eloc, void *stream);

/* O[B][ld] = d ln|psi| / d theta in the layout above (o_dtype 0 f64, 1 f32),
 * with the environments of the exact contraction (and the MLP backward).
 * Replaces: compute_grads_gpu for fPEPS / NN-fTNS (GPU/vmc_utils.py:1389-1634). */
int vmcx_peps_logderiv(const vmcx_peps_model *m, int64_t B, const int64_t *configs,
                       void *workspace, int64_t ws_bytes, int64_t chunk, void *O, int o_dtype,
                       int64_t ld, double *sign, double *logabs, void *stream);

/* The factors of the per-sample log-derivative instead of O (NN-fTNS:
 * row s of F, [B][ld] float64, ld >= n_params + 2N + 4 hidden, is
 * [g | n | dz1 | h1 | dz2 | h2] with g = d ln|psi| / d theta_eff; plain fPEPS
 * (hidden 0): g alone, ld >= n_params).  O = [g, dz1 n^T, dz1, dz2 h1^T, dz2,
 * eta g h2^T, eta g], the blocks vmcx_sr_minsr_factored consumes.  Mode 1
 * (multiplicative): row s is [g (1 + eta y) | n | dz1 | h1 | dz2 | h2 |
 * g theta] (ld >= 2 n_params + 2N + 4 hidden) and O = [g (1 + eta y), dz1 n^T,
 * dz1, dz2 h1^T, dz2, eta (g theta) h2^T, eta g theta].  Also
 * writes sign / logabs.  Replaces: compute_grads_gpu for NN-fTNS
 * (GPU/vmc_utils.py:1389-1634, GPU/models/NNfTNS.py:280-955) when the SR
 * step takes the factored form. */
int vmcx_peps_logderiv_factors(const vmcx_peps_model *m, int64_t B, const int64_t *configs,
                               void *workspace, int64_t ws_bytes, int64_t chunk, double *F,
                               int64_t ld, double *sign, double *logabs, void *stream);

/*
 * Learned Metropolis proposal (no counterpart in the reference; the closest is
 * the external-proposal sampler MetropolisMPSSamplerSpinful, sampler.py:
 * 2120-2238, acceptance min(1, P'/P * q/q') at :2214-2218).
 *   Net: occupation grid (n_up, n_dn) of an Lx x Ly lattice -> three 3x3
 *   convolutions (zero padding, C channels, ReLU) -> per-edge head
 *   logit[e][t] = bA[t] + A[t][:C] . h[:, i] + A[t][C:] . h[:, j].
 *   params: FP32 flat [W1 (C,2,3,3) | b1 | W2 (C,C,3,3) | b2 | W3 | b3 |
 *   A (3, 2C) | bA (3)], vmcx_lp_n_params(C) entries, C <= 32, N <= 64.
 *   Moves m = 3 e + t, edges e = (i, j), i < j, sorted ((s, s+1) then
 *   (s, s+Ly)); t = 0 hop of an up electron, 1 hop of a down electron,
 *   2 exchange of a down and an up electron; q(m|x) = softmax over the valid
 *   moves; every move is its own reverse.
 *   Draws: Philox4x32-10, counter (step lo, step hi, walker, 0x4C50), key =
 *   seed; u_choice = top 53 bits of (o0:o1), u_acc of (o2:o3); injected
 *   arrays (may be NULL) replace them.
 */
int vmcx_lp_n_params(int C);
/* 3 x number of edges of the Lx x Ly open lattice; -1 on bad sizes. */
int vmcx_lp_n_moves(int Lx, int Ly);
/* logits [B][n_moves] float64, -inf for invalid moves.  precision 0: FP32
 * convolutions on the vector ALUs (any C); 1: layer 1 FP32, layers 2-3 on
 * the BF16 matrix cores with FP32 accumulation and BF16 activations (C = 32);
 * the head accumulates in FP64 either way. */
int vmcx_lp_logits(const float *params, int C, int Lx, int Ly, int64_t B,
                   const int64_t *configs, double *logits, int precision, void *stream);
/* Draw m ~ q(.|x) for every walker, write the proposed configuration and
 * ln q(m|x), then evaluate the net at x' for ln q(m|x') (move = -1 when x
 * has no valid move).  logits_ws: [B][n_moves] float64 scratch; precision as
 * for vmcx_lp_logits (the same net at x and x', so the chain stays exact). */
int vmcx_lp_propose(const float *params, int C, int Lx, int Ly, int64_t B,
                    const int64_t *configs, uint64_t seed, int64_t step, const double *u_choice,
                    double *logits_ws, int32_t *move, double *logq_fwd, int64_t *proposed,
                    double *logq_rev, int precision, void *stream);
/* vmcx_lp_propose with the logits at the current configurations given
 * (logits_x, [B][n_moves]): the net runs once, on the proposed
 * configurations (logits_p out).  After the accept step,
 * vmcx_lp_commit_logits copies the accepted walkers' rows of logits_p into
 * logits_x, so a chain evaluates the net once per proposal; the draws and
 * decisions are those of vmcx_lp_propose (the same net on the same
 * configuration gives the same bits).  No counterpart in the reference. */
int vmcx_lp_propose_given(const float *params, int C, int Lx, int Ly, int64_t B,
                          const int64_t *configs, const double *logits_x, uint64_t seed,
                          int64_t step, const double *u_choice, double *logits_p,
                          int32_t *move, double *logq_fwd, int64_t *proposed, double *logq_rev,
                          int precision, const int64_t *step_dev, void *stream);
int vmcx_lp_commit_logits(int64_t B, int n_moves, const uint8_t *accepted,
                          const double *logits_p, double *logits_x, void *stream);
/* Metropolis-Hastings: accept when u_acc < |psi'|^2/|psi|^2 q(m|x')/q(m|x),
 * from (psign, plogabs) of the proposed configurations; commits configs,
 * sign, logabs, n_accept (may be NULL) and accepted[B] (may be NULL).
 * step_dev (may be NULL; also in vmcx_lp_propose_given): a device int64
 * added to step when the kernels run (draws advanced by vmcx_counter_add
 * inside a captured graph). */
int vmcx_lp_accept(int64_t B, int n_sites, uint64_t seed, int64_t step, const double *u_acc,
                   const int32_t *move, const int64_t *proposed, const double *psign,
                   const double *plogabs, const double *logq_fwd, const double *logq_rev,
                   int64_t *configs, double *sign, double *logabs, int64_t *n_accept,
                   uint8_t *accepted, const int64_t *step_dev, void *stream);
/* vmcx_lp_accept, vmcx_lp_commit_logits and (z1 != NULL) vmcx_masked_rows of
 * the walkers' state rows (z1p -> z1, row_bytes a multiple of 4) in one
 * launch, one wave per walker: the same decisions and results bit for bit.
 * Replaces: the accept / commit step of the learned-proposal chain
 * (GPU/sampler.py:216-388 with the MH rule of sampler.py:2214-2218). */
int vmcx_lp_accept_commit(int64_t B, int n_sites, uint64_t seed, int64_t step,
                          const double *u_acc, const int32_t *move, const int64_t *proposed,
                          const double *psign, const double *plogabs, const double *logq_fwd,
                          const double *logq_rev, int64_t *configs, double *sign, double *logabs,
                          int64_t *n_accept, uint8_t *accepted, const int64_t *step_dev,
                          int n_moves, const double *logits_p, double *logits_x,
                          int64_t row_bytes, const void *z1p, void *z1, void *stream);
/* grad[n_params] = sum_b d(sum_m dlogits[b][m] logits[b][m]) / d params
 * (FP32; per-configuration rows in `workspace`, summed in a fixed order). */
int64_t vmcx_lp_backward_workspace_bytes(int C, int64_t B);
int vmcx_lp_backward(const float *params, int C, int Lx, int Ly, int64_t B,
                     const int64_t *configs, const double *dlogits, void *workspace,
                     int64_t ws_bytes, float *grad, void *stream);
/* out[B][n_moves][N]: every move applied to each configuration (a copy for
 * invalid moves): the candidates whose amplitudes give the heat-bath target. */
int vmcx_lp_candidates(int64_t B, int Lx, int Ly, const int64_t *configs, int64_t *out,
                       void *stream);
/* y += a x (FP32): the parameter update of the proposal net. */
int vmcx_lp_axpy(int64_t n, float a, const float *x, float *y, void *stream);

/*
 * Point-group images for the symmetry-projected wavefunction (replaces the
 * x[:, perms] gather and compute_fermion_signs of GPU/models/symmetry.py:
 * 660-770, 893-951).  perms[g][s] (int32, G x N): site whose occupation
 * image g puts on site s, images[b][g][s] = configs[b][perms[g][s]];
 * inv_perms: the inverse permutations.  signs[b][g] = (-1)^(inversions,
 * among the occupied modes of configs[b] in the interleaved order
 * [d0, u0, d1, u1, ...], of the mode permutation 2s + spin ->
 * 2 inv_perms[g][s] + spin).  N <= 64, G <= 64.
 */
int vmcx_sym_images(int64_t B, int n_sites, int n_group, const int64_t *configs,
                    const int32_t *perms, const int32_t *inv_perms, int64_t *images,
                    double *signs, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* VMCX_H */
