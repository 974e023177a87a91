/* numpyro_amd: MI355X-native HMC/NUTS engine, C ABI.
 *
 * Plain pointers and sizes only.  Every pointer named `d_*` is a device
 * pointer on the current HIP device; all other pointers are host memory.
 * Every call returns 0 on success or a negative NAX_E* code; the message is
 * available from nax_last_error().  Numerical events (NaN energies,
 * divergences) are never errors: they become divergence flags, as in
 * numpyro/infer/hmc_util.py:871-874.
 *
 * Each entry point cites the reference interface it replaces (numpyro
 * 0.21.0, paths relative to the reference repository root).
 */
#ifndef NUMPYRO_AMD_H
#define NUMPYRO_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NAX_OK 0
#define NAX_EINVAL (-22)
#define NAX_ENOMEM (-12)
#define NAX_EHIP (-5)
#define NAX_ESTATE (-71)

#define NAX_ALGO_NUTS 0
#define NAX_ALGO_HMC 1

#define NAX_X_F32 0
#define NAX_X_BF16 1
#define NAX_X_FP8 2 /* OCP e4m3fn + power-of-two column scales, VALU tile kernel */

typedef struct nax_sampler nax_sampler;

/* Kernel configuration: the keyword arguments of hmc()'s init_kernel
 * (numpyro/infer/hmc.py:193-212) and of NUTS/HMC (hmc.py:625-686, 916-951).
 * A non-positive trajectory_length / num_steps means None. */
typedef struct {
  int32_t num_chains;
  int32_t dim;
  int32_t algo;
  int32_t num_warmup;
  int32_t max_tree_depth_warmup;
  int32_t max_tree_depth;
  int32_t adapt_step_size;
  int32_t adapt_mass_matrix;
  int32_t regularize_mass_matrix;
  int32_t partitionable_keys; /* jax_threefry_partitionable (jax>=0.5 default: 1) */
  int32_t num_steps;
  /* find_heuristic_step_size (hmc.py:206, 319-323): run
   * find_reasonable_step_size (hmc_util.py:314-384) at init and at the end
   * of every middle adaptation window (hmc_util.py:572-575, 618-623). */
  int32_t find_heuristic_step_size;
  double step_size;
  double target_accept_prob;
  double trajectory_length;
  double max_delta_energy; /* hmc.py:188: 1000 */
  /* dense_mass=True (hmc.py:201): full inverse mass matrix per chain, dense
   * Welford adaptation (hmc_util.py:133-239); dim <= 1024. */
  int32_t dense_mass;
  int32_t reserved;
} nax_config;

/* Per-iteration collection buffers, fori_collect semantics
 * (numpyro/util.py:321-454).  Any pointer may be NULL (not collected).
 * Shapes: z [C, size, dim] fp32; the others [C, size]. */
typedef struct {
  int64_t start_idx;
  int32_t thinning;
  int32_t size;
  float* d_z;
  double* d_potential_energy;
  double* d_energy;
  float* d_accept_prob;
  float* d_mean_accept_prob;
  int32_t* d_num_steps;
  uint8_t* d_diverging;
  float* d_step_size;
} nax_collect;

/* Host view of the per-chain state (HMCState, hmc.py:31-48, with the
 * HMCAdaptState of hmc_util.py:16-30).  Device arrays are [C, dim] fp32
 * unless stated (inverse mass matrix and mass_matrix_sqrt: [C, dim, dim]
 * under dense_mass); host scalars are [C]. */
typedef struct {
  float* d_z;
  float* d_z_grad;
  float* d_inverse_mass_matrix;
  float* d_mass_matrix_sqrt;
  int32_t* i;
  double* potential_energy;
  double* energy;
  double* accept_prob;
  double* mean_accept_prob;
  int32_t* num_steps;
  int32_t* diverging;
  double* step_size;
  uint32_t* rng_key; /* [C, 2] */
} nax_state_view;

/* Host view of the rest of HMCAdaptState (hmc_util.py:16-30): the
 * dual-averaging state ss_state = (x_t, x_avg, g_avg, t, prox_center)
 * (dual_averaging, hmc_util.py:97-130), the Welford state mm_state =
 * (mean, m2, n) (welford_covariance, hmc_util.py:160-196), window_idx and
 * the adapter's rng_key (hmc_util.py:565, 655-656).  Host arrays are [C];
 * d_mm_mean is [C, dim] fp64 and d_mm_m2 [C, dim] fp64, or [C, dim, dim]
 * fp64 under dense_mass (m2[i][j] = sum of delta_post_i * delta_pre_j).
 * Any pointer may be NULL (that part is not copied). */
typedef struct {
  double* ss_x_t;
  double* ss_x_avg;
  double* ss_g_avg;
  int32_t* ss_t;
  double* ss_prox_center;
  double* d_mm_mean;
  double* d_mm_m2;
  int32_t* mm_n;
  int32_t* window_idx;
  uint32_t* rng_key; /* [C, 2] */
} nax_adapt_view;

const char* nax_last_error(void);

/* Replaces hmc(potential_fn, algo=...) (hmc.py:113) + init_kernel's static
 * configuration (hmc.py:193-212). */
int nax_create(const nax_config* cfg, nax_sampler** out);
void nax_destroy(nax_sampler* h);

/* Built-in potential: Bayesian logistic regression, the model of
 * examples/covtype.py:66-71 (coefs ~ Normal(0, prior_scale), obs ~
 * Bernoulli(logits=X @ coefs)), replacing potential_energy
 * (numpyro/infer/util.py:333-358) + jax.value_and_grad (hmc_util.py:242-252).
 * d_x: [n, d] row-major fp32; d_y: [n] fp32 in {0,1}.  x_dtype selects the
 * storage the kernels read (fp32, or bf16 on the matrix cores); the library
 * keeps its own converted copy for bf16. */
int nax_model_logistic(nax_sampler* h, const float* d_x, const float* d_y, int64_t n,
                       int32_t d, double prior_scale, int32_t x_dtype);

/* Built-in potential: the stochastic-volatility model of
 * examples/stochastic_volatility.py:58-66 (sigma ~ Exponential(50), s ~
 * GaussianRandomWalk(sigma, T), nu ~ Exponential(0.1), r ~ StudentT(nu, 0,
 * exp(s))), with its hand-written adjoint.  d_returns: [T] fp32; the sampler
 * dim must be T + 2.  noncentred = 0: latent [log nu, s (T), log sigma] (the
 * reference's model, sorted site names); 1: latent [log nu, log sigma, zeta
 * (T)] with s = sigma * cumsum(zeta).  Sampling runs as one persistent kernel
 * per call (one wavefront per chain). */
int nax_model_sv(nax_sampler* h, const float* d_returns, int32_t T, int32_t noncentred);

/* nax_model_sv with the two prior rates as arguments: sigma ~
 * Exponential(sigma_rate), nu ~ Exponential(nu_rate) (both finite and > 0,
 * else NAX_EINVAL).  The target of the model compiler for stochastic-
 * volatility models written with other rates (the reference builds the same
 * potential from the model function, numpyro/infer/util.py:300-358). */
int nax_model_sv_ex(nax_sampler* h, const float* d_returns, int32_t T, int32_t noncentred,
                    double sigma_rate, double nu_rate);

/* Built-in potential: `groups` independent non-centred hierarchical logistic
 * regressions (eta ~ Normal(0,1)^D_i, mu ~ Normal(0,1), tau ~ HalfCauchy(1),
 * beta = mu + tau eta, y ~ Bernoulli(logits = X_i beta)), chains_per_group
 * chains each (num_chains = groups * chains_per_group, chain c belongs to
 * group c / chains_per_group).  Ragged data: d_x holds X_i row-major at
 * element offset d_elt_off[i] (int64 [groups]); d_y holds y_i at row offset
 * d_row_off[i] (int64 [groups + 1], N_i = d_row_off[i+1] - d_row_off[i]);
 * d_dims (int32 [groups]) = D_i <= 64.  Chain latent (sorted site names):
 * [eta (D_i), mu, log tau]; the sampler dim must be max D_i + 2 and each
 * chain uses its own D_i + 2.  Sampling runs as one persistent launch, a
 * workgroup per model. */
int nax_model_hier(nax_sampler* h, const float* d_x, const float* d_y, const int64_t* d_row_off,
                   const int64_t* d_elt_off, const int32_t* d_dims, int32_t groups,
                   int32_t chains_per_group);

/* nax_model_logistic with the kernel choice for fp32 storage: 0 = auto
 * (one wavefront per chain in a persistent launch for <= 16 chains on
 * n*d <= 2^20, d <= 64; otherwise the VALU tile kernel or, from 32 chains
 * up, the fp32 MFMA kernel, whichever the autotuner times faster), 1 = VALU
 * tile kernel, 2 = fp32 MFMA, 3 = one wavefront per chain.  bf16 storage
 * always uses the matrix cores: 0 (auto) or 4 keep one bf16 copy of X and
 * read it transposed for X^T R with bf16 residuals; 2 keeps a second, fp16
 * copy for X^T R with fp16 residuals (twice the X bytes per evaluation, a
 * gradient rounding of 2^-11 instead of 2^-8 per term, and max|X| < 60000).
 * fp8 storage (NAX_X_FP8; the library quantises d_x as nax_fp8_quantize
 * does and keeps the codes) runs on the VALU tile kernel: 0 (auto, row
 * splits autotuned) or 1.  Other kernel values with bf16 or fp8 storage,
 * and 4 with fp32 storage, are refused (NAX_EINVAL). */
int nax_model_logistic_ex(nax_sampler* h, const float* d_x, const float* d_y, int64_t n,
                          int32_t d, double prior_scale, int32_t x_dtype, int32_t kernel);

/* Subsampled logistic regression for HMCECS (numpyro/infer/hmc_gibbs.py:502-,
 * the covtype model of examples/covtype.py:66-71 with subsample_size, run as
 * in examples/covtype.py:154-165).  Each chain's potential uses its own
 * subsample of subsample_size rows, d_idx [C, subsample_size] int32 (caller-
 * owned, initialised by the caller, updated by nax_ecs_gibbs).  proxy_degree
 * 0: the plate's n/m scaling of the subsample log-likelihood; 1 or 2: the
 * difference estimator with a Taylor proxy around d_ref [d] fp64
 * (numpyro/contrib/ecs_proxies.py:23-50, 95-250).  d_keys [C, 2] are the
 * HMCECS state keys (HMCECSState.rng_key), advanced by nax_ecs_gibbs.
 * subsample_size <= 4096, dim <= 512. */
int nax_model_logistic_ecs(nax_sampler* h, const float* d_x, const float* d_y, int64_t n,
                           int32_t d, double prior_scale, int32_t subsample_size,
                           int32_t num_blocks, int32_t proxy_degree, const double* d_ref,
                           int32_t* d_idx, uint32_t* d_keys);

/* One HMCECS Gibbs update of every chain's subsample (hmc_gibbs.py:640-672,
 * block update ecs_proxies.py:58-92): a block of indices is redrawn, U and
 * dU at the chain's current position are evaluated with the proposal, and
 * the proposal is accepted with probability min(1, exp(U - U_new)).  Call it
 * before each nax_run of one iteration.  d_accept_prob [C] fp32 or NULL. */
int nax_ecs_gibbs(nax_sampler* h, float* d_accept_prob);

/* Potential supplied by the caller (a potential_fn, hmc.py:127-130): the
 * engine asks for U and dU at d_z_request, the caller writes them into the
 * gradient buffers, see nax_advance. */
int nax_model_external(nax_sampler* h);

/* Initial state: init_kernel (hmc.py:193-362) for C chains.
 * d_z0 [C, dim]; keys [C, 2] host (the per-chain rng_key of init_kernel);
 * d_inverse_mass_matrix [dim] ([dim, dim] under dense_mass) or NULL
 * (identity), shared by all chains. */
int nax_init(nax_sampler* h, const float* d_z0, const uint32_t* keys,
             const float* d_inverse_mass_matrix);

/* Runs sample_kernel (hmc.py:459-530) on every chain until each has done
 * `target_iter` iterations in total, collecting per `col` (may be NULL).
 * Chains run their iterations asynchronously between gradient evaluations.
 * Built-in potentials only.  *grad_evals (may be NULL) receives the number of
 * useful gradient evaluations summed over chains. */
int nax_run(nax_sampler* h, int64_t target_iter, const nax_collect* col, int64_t* grad_evals);

/* One state-machine step for an external potential: consumes the
 * gradients in the buffers of nax_buffers(), advances each chain to its next
 * gradient request (or to target_iter), and returns in *pending the number
 * of chains that requested a gradient (0: done). */
int nax_advance(nax_sampler* h, int64_t target_iter, const nax_collect* col, int32_t* pending);

/* Device buffers of the gradient exchange: z requests [C, ld] fp32,
 * potential energies [C] fp64, gradients [C, ld] fp32. */
int nax_buffers(nax_sampler* h, float** d_z_request, double** d_potential_energy,
                float** d_grad, int32_t* ld);

/* Copies the requested positions out to d_z_out [C, dim] and/or the
 * caller's U [C] fp64 and dU [C, dim] fp32 in (any pointer may be NULL);
 * the dense-layout form of nax_buffers. */
int nax_exchange(nax_sampler* h, float* d_z_out, const double* d_pe_in, const float* d_grad_in);

/* The bijection that unconstrains the model's positive latent sites
 * (potential_energy's _unconstrain_reparam, numpyro/infer/util.py:300-330):
 * kind 0 = ExpTransform, biject_to(constraints.positive) (transforms.py:
 * 606-652, 2115-2118, the default); kind 1 = SoftplusTransform,
 * biject_to(constraints.softplus_positive) (transforms.py:1156-1175,
 * 2213-2215).  Applies to sigma and nu of the stochastic-volatility model
 * and tau of the hierarchical model; NAX_EINVAL for a model without
 * positive sites, NAX_ESTATE after nax_init. */
int nax_set_positive_transform(nax_sampler* h, int32_t kind);

/* The launch shapes the model's autotuner picked (waves per chain of a
 * persistent launch, row splits / kernel form of the logistic kernels,
 * chains per workgroup and waves per chain of the hierarchical kernel, with
 * the timings it compared), as text
 * into buf (at most len bytes with the terminating zero; empty before the
 * first tuned launch).  No counterpart in the reference, which compiles one
 * XLA program per model. */
int nax_tuning(nax_sampler* h, char* buf, int32_t len);

/* Fixes the launch shape the autotuner would pick: waves per chain (1, 2,
 * 4, 8, 16) of the per-chain persistent models (stochastic volatility, the
 * one-wave-per-chain logistic form), or, for the hierarchical model, chains
 * per workgroup kb (a divisor of chains_per_group) and waves per chain
 * r (1, 2 or 4; kb r <= 16) as shape = kb + 32 (r - 1).  Draws depend on
 * the waves per chain (the order of its sums), not on the chains per
 * workgroup; a fixed shape makes runs repeat bit for bit.  A fixed shape
 * also sets the waves per chain of nax_potential_grad for these models
 * (one wave per chain otherwise).  No counterpart in the reference. */
int nax_set_launch_shape(nax_sampler* h, int32_t shape);

/* HMCState / post_warmup_state access (mcmc.py:558-587): copies the
 * per-chain state out of / into the sampler. */
int nax_get_state(nax_sampler* h, const nax_state_view* v);
int nax_set_state(nax_sampler* h, const nax_state_view* v);

/* HMCAdaptState's ss_state / mm_state / window_idx / rng_key out of / into
 * the sampler, so that a state taken mid-warmup resumes exactly (the
 * reference's functional state, hmc.py:459-530 and mcmc.py:558-587, carries
 * them in hmc_state.adapt_state). */
int nax_get_adapt_state(nax_sampler* h, const nax_adapt_view* v);
int nax_set_adapt_state(nax_sampler* h, const nax_adapt_view* v);

/* The FP8 storage of nax_model_logistic(_ex) with x_dtype NAX_X_FP8 (the
 * fp8 variant of the C5 precision configuration; the reference computes
 * jnp.dot(data, coefs), examples/covtype.py:70, in the data's dtype): per
 * column d the power-of-two scale s_d >= max_n |x[n][d]| / 448 (1 for an
 * all-zero column) into d_scales [d], and the round-to-nearest-even OCP
 * e4m3fn byte of x[n][d] / s_d into d_codes [n][d].  The model's design is
 * then x_q = e4m3(code) * s_d, exact in fp32. */
int nax_fp8_quantize(const float* d_x, int64_t n, int32_t d, uint8_t* d_codes, float* d_scales);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the potential and gradient of the logistic model
 * (coefs ~ Normal(0, prior_scale), obs ~ Bernoulli(logits = X @ coefs)) at
 * d_z [count, ld] fp32 -> d_pe [count] fp64, d_grad [count, ld] fp32, with
 * the kernels of nax_potential_grad on a logistic model in the fp32 tile
 * form (nax_model_logistic_ex kernel 1), without a sampler handle.  d_x
 * [n, d] fp32 row-major, d_y [n] fp32; d <= 320.  Allocates nothing: the
 * row-split partials live in a 16 MiB device array of the library (one per
 * GPU, held until the process ends); count * (8 + 4 d) bytes must fit it
 * (else NAX_EINVAL).  Asynchronous on `stream` (a hipStream_t; NULL = the
 * null stream).  Calls that overlap on one GPU share that array and race. */
int nax_logistic_potential_grad_noalloc(const float* d_x, const float* d_y, int64_t n, int32_t d,
                                        double prior_scale, const float* d_z, int32_t ld,
                                        int32_t count, double* d_pe, float* d_grad,
                                        void* stream);

/* Potential and gradient of the built-in model at d_z [count, ld]:
 * potential_energy + value_and_grad (util.py:333-358, hmc_util.py:242-252). */
int nax_potential_grad(nax_sampler* h, const float* d_z, int32_t ld, int32_t count,
                       double* d_pe, float* d_grad);

/* L velocity_verlet steps (hmc_util.py:262-311) for C chains in place:
 * d_z, d_r, d_grad [C, dim]; d_pe [C]; d_inverse_mass_matrix [dim]. */
int nax_leapfrog(nax_sampler* h, float* d_z, float* d_r, float* d_grad, double* d_pe,
                 double step_size, int32_t num_steps, const float* d_inverse_mass_matrix);

/* jax.random primitives (the call sites listed in hmc.py:102, 335, 472 and
 * hmc_util.py:804, 1162): threefry2x32-20 over counters; random.split;
 * random.normal(key, (n,)). keys/counters are device arrays. */
int nax_threefry2x32(const uint32_t* d_key /*2*/, const uint32_t* d_x0, const uint32_t* d_x1,
                     uint32_t* d_out0, uint32_t* d_out1, int64_t n);
int nax_random_split(const uint32_t* d_key /*2*/, int32_t num, int32_t partitionable,
                     uint32_t* d_out /*[num,2]*/);
/* random.randint(keys[c], (n,), lo, hi) for nk keys (jax.random.randint's
 * two-draw construction, as ecs_proxies.py:58-71 consumes it): d_out [nk, n]. */
int nax_randint(const uint32_t* d_keys /*[nk,2]*/, int32_t nk, int32_t n, int32_t lo, int32_t hi,
                int32_t partitionable, int32_t* d_out);
/* random.split(keys[c], num) for nk keys: d_out [nk, num, 2]. */
int nax_split_keys(const uint32_t* d_keys /*[nk,2]*/, int32_t nk, int32_t num,
                   int32_t partitionable, uint32_t* d_out);
int nax_random_normal(const uint32_t* d_key /*2*/, int32_t n, int32_t partitionable,
                      float* d_out);

/* Per-chain init_kernel keys the way MCMC.run + HMC.init derive them from
 * one key: split(key, C)[c] (mcmc.py:670-671), then split(.)[0]
 * (hmc.py:744-750).  d_out [C, 2]. */
int nax_chain_keys(const uint32_t* d_key /*2*/, int32_t num_chains, int32_t partitionable,
                   uint32_t* d_out);

/* Progress report from nax_run's host loop: cb(user, ticks, pending) after
 * every `every_ticks` gradient ticks (rounded up to the loop's batch of 16);
 * pending = chains still running.  cb = NULL switches it off. */
int nax_set_progress(nax_sampler* h, void (*cb)(void* user, int64_t ticks, int32_t pending),
                     void* user, int64_t every_ticks);

/* Timing of the dominant (gradient) kernel inside nax_run, measured with HIP
 * events on the sampler's stream: total milliseconds and launches since the
 * last reset. */
int nax_kernel_timing(nax_sampler* h, int32_t enable, double* total_ms, int64_t* launches);

/* Launch mode of nax_run's tick loop (no counterpart in the reference, whose
 * loop is one XLA program): with enable != 0 (the default) batches of ticks
 * are captured once per call as a hipGraph and replayed; 0 enqueues every
 * kernel directly.  Results are identical either way. */
int nax_set_graphs(nax_sampler* h, int32_t enable);

/* Diagnostics of collected draws d_x [chains, draws, params] fp32 (device):
 * effective_sample_size (numpyro/diagnostics.py:158-203, biased
 * autocovariance, Geyer's initial monotone sequence) into d_ess [params] and
 * split_gelman_rubin (diagnostics.py:64-80) into d_split_rhat [params] (may
 * be NULL), both fp64 on the device.  draws <= 4096.  Synchronous. */
int nax_diagnostics(const float* d_x, int32_t chains, int32_t draws, int32_t params,
                    double* d_ess, double* d_split_rhat);

#ifdef __cplusplus
}
#endif

#endif
