// Host-side internals shared by the translation units of libnumpyro_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/numpyro_amd.h"

namespace nax {

constexpr int kMaxWindows = 64;

// Engine configuration passed by value to the kernels.
struct EngineCfg {
  int C, D, Dp, nvec;
  int algo;  // 0 NUTS, 1 HMC
  int num_warmup;
  int max_depth_warmup, max_depth_post, max_depth_alloc;
  int adapt_step_size, adapt_mass_matrix, regularize;
  int partitionable;
  int find_heuristic;         // find_reasonable_step_size at init / window ends
  int fixed_num_steps;        // <= 0: None
  double trajectory_length;   // <= 0: None
  double target_accept, max_delta_energy;
  int num_windows;
  int window_end[kMaxWindows];
  // Per-chain latent size (ragged models), nullptr when every chain has D.
  // Vectors are strided by Dp; entries past a chain's own size stay zero.
  const int* chain_dim;
  // Dense mass matrix (dense_mass=True): per chain [Dp][Dp] blocks of the
  // transposed inverse mass matrix, the transposed mass_matrix_sqrt and the
  // fp64 Welford co-moment matrix (also the factorisation workspace).
  int dense;
  float* dimm_t;
  float* dmsq_t;
  double* dm2;
};

// Per-chain scalar state (one struct per chain in device memory).  All
// fields are wave-uniform: one wavefront owns one chain.
struct ChainState {
  // HMCState (numpyro/infer/hmc.py:31-48)
  int32_t i, phase;
  uint32_t rng_key[2];
  double pe, energy, accept_prob, mean_accept_prob;
  int32_t num_steps, diverging;
  // build_tree loop (numpyro/infer/hmc_util.py:1088-1180)
  uint32_t tree_key[2], sub_key[2], top_key[2], leaf_key[2];
  int32_t going_right, depth, max_depth_cur, turning;
  int32_t t_diverging, num_proposals, s_num, s_turning;
  int32_t s_diverging, hmc_step, hmc_num_steps, pad0;
  double energy_current, weight, sum_accept, peP, energyP;
  double s_weight, s_sum_accept, s_peP, s_energyP;
  double pe_edge, hmc_eps;
  // HMCAdaptState (numpyro/infer/hmc_util.py:16-30)
  double step_size, da_x, da_xavg, da_gavg, da_prox;
  int32_t da_t, window_idx, welford_n, pad1;
  uint32_t adapt_key[2];
  uint32_t trans_key[2];
  // find_reasonable_step_size loop (hmc_util.py:314-384)
  uint32_t fs_key[2];
  int32_t fs_dir, fs_last_dir, fs_ret, pad2;
  double fs_step, fs_e0;
};

// Collection of per-iteration fields, fori_collect semantics
// (numpyro/util.py:321-454): the state after loop index `it` goes to slot
// (it - start_idx) / thinning, the last of each group of `thinning` wins.
struct Collect {
  int64_t start_idx;
  int32_t thinning, size;
  float* z;
  double* potential_energy;
  double* energy;
  float* accept_prob;
  float* mean_accept_prob;
  int32_t* num_steps;
  uint8_t* diverging;
  float* step_size;
};

// A potential U(z) and its gradient for C chains at once.
struct Potential {
  virtual ~Potential() {}
  virtual int dim() const = 0;
  // z: [C, ldz] fp32; writes pe[C] (fp64) and g: [C, ldg] fp32.
  virtual hipError_t eval(hipStream_t s, const float* z, int ldz, int C, double* pe,
                          float* g, int ldg) = 0;
  // The same restricted to the chains act[0..*nact) (device memory, written
  // by the tick kernel); a potential may evaluate more chains than asked.
  virtual hipError_t eval_active(hipStream_t s, const float* z, int ldz, int C, const int* act,
                                 const int* nact, double* pe, float* g, int ldg) {
    return eval(s, z, ldz, C, pe, g, ldg);
  }
  // Bytes / flops one eval of C chains moves / does (algorithmic figures).
  virtual double alg_bytes(int C) const { return 0; }
  virtual double alg_flops(int C) const { return 0; }
  virtual const char* name() const { return "potential"; }
  // Launch shapes the potential's autotuner chose (empty when it has none).
  std::string tuning;
  // Fixes the tuned launch shape instead of timing candidates (waves per
  // chain of per-chain persistent models, chains per workgroup of the
  // hierarchical model); false when the model has no such shape or the
  // value is not one of its candidates.
  virtual bool force_launch_shape(int shape) { return false; }
  // Positive latent sites by ExpTransform (kind 0) or SoftplusTransform
  // (kind 1); false for a model without positive sites.
  virtual bool set_positive_transform(int kind) { return false; }
  // Latent size of each chain on the device ([C] int32) for ragged models.
  virtual const int* chain_dims() const { return nullptr; }
  // Models whose potential needs no data shared across chains run the whole
  // sampling loop in one persistent launch per call.
  virtual bool persistent() const { return false; }
  // d_target: device int64, the iteration every chain runs to.  A launch
  // enqueues work only (no allocation, no synchronisation), so that it can
  // be captured in a graph; prepare_persistent runs first, uncaptured
  // (autotuning, workspaces).
  virtual hipError_t run_persistent(hipStream_t s, const EngineCfg& cfg, const Collect& col,
                                    const int64_t* d_target, ChainState* st, float* vec,
                                    double* wel, float* zreq, float* gout, double* peout,
                                    int64_t* gevals) {
    return hipErrorNotSupported;
  }
  virtual hipError_t prepare_persistent(hipStream_t s, const EngineCfg& cfg, ChainState* st,
                                        float* vec, double* wel, float* zreq, float* gout,
                                        double* peout) {
    return hipSuccess;
  }
  // A model whose launch shape is best chosen on trees of working size sets
  // tune_at > 0: chains run its default shape below that iteration, the
  // sampler ends a chunk exactly there and calls prepare_persistent again
  // (chain_iter holds the chains' iteration at each call), so every run of
  // the same keys uses the same shape at the same iterations.
  int64_t tune_at = 0;
  int64_t chain_iter = 0;
  // HMCECS models: one Gibbs update of every chain's subsample at its
  // current position (z, stride ldv), accepting into st[c].pe and the state
  // gradient gst (stride ldv); accept[c] receives the acceptance probability.
  virtual hipError_t ecs_gibbs(hipStream_t s, ChainState* st, const float* z, float* gst, int ldv,
                               float* accept) {
    return hipErrorNotSupported;
  }
  // When set, the potential records these events on its stream immediately
  // before and after its main (dominant) kernel.
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  void mark_begin(hipStream_t s) { if (ev_begin) hipEventRecord(ev_begin, s); }
  void mark_end(hipStream_t s) { if (ev_end) hipEventRecord(ev_end, s); }
};

Potential* make_logistic(const float* x, const float* y, int64_t n, int d, double prior_scale,
                         int x_dtype, int fuse_variant, int num_chains, hipStream_t s,
                         std::string* err);

Potential* make_logistic_chain(const float* x, const float* y, int64_t n, int d,
                               double prior_scale, hipStream_t s, std::string* err);

Potential* make_hier(const float* x, const float* y, const int64_t* row_off, const int64_t* elt_off,
                     const int* dims, int groups, int chains_per_group, int num_chains, int Dp,
                     hipStream_t s, std::string* err);
Potential* make_sv(const float* returns, int T, int noncentred, double sigma_rate, double nu_rate,
                   hipStream_t s, std::string* err);

Potential* make_ecs_logistic(const float* x, const float* y, int64_t n, int d, double prior_scale,
                             int m, int num_blocks, int proxy, const double* ref, int* idx,
                             uint32_t* keys, int num_chains, int Dp, bool partitionable,
                             hipStream_t s, std::string* err);

void set_error(const std::string& msg);

}  // namespace nax
