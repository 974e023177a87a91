// NUTS / HMC engine: per-chain state machine kernels and the C ABI host driver.
//
// One wavefront owns one chain.  The `advance` kernel runs each chain's
// control flow -- sample_kernel (numpyro/infer/hmc.py:459-530), build_tree
// (numpyro/infer/hmc_util.py:1088-1180) with the iterative subtree, the
// checkpointed U-turn checks and the uniform/biased transitions
// (hmc_util.py:710-1085), and the warmup adapter (hmc_util.py:518-707) --
// until the chain needs its next gradient.  The positions of all chains are
// then handed to one batched potential+gradient evaluation, so the X.B and
// X^T.R contractions of the logistic likelihood run as GEMMs over chains.
// Chains do not wait for each other: a chain that ends its trajectory starts
// its next iteration immediately.  Each chain is still exactly the Markov
// chain of the reference (same key consumption, same decisions); only the
// interleaving of chains in time differs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "nax_device.h"
#include "nax_internal.h"

namespace nax {

// ---- per-chain vector slots (each [Dp] fp32) --------------------------------
enum {
  V_Z = 0, V_G,                    // HMCState.z, z_grad
  V_ZL, V_RL, V_GL, V_ZR, V_RR, V_GR,  // tree edges (the subtree's outer edge lives here too)
  V_ZP, V_GP,                      // tree proposal
  V_RSUM,                          // tree r_sum
  V_SZP, V_SGP, V_SRSUM,           // subtree proposal and r_sum
  V_IMM, V_MSQRT,                  // inverse mass matrix diag, mass_matrix_sqrt diag
  V_R0,                            // momentum drawn by init_kernel
  V_CK                             // r_ckpts[depth], then r_sum_ckpts[depth]
};

enum {
  PH_INIT_WAIT = 0, PH_ITER_START, PH_DOUBLING, PH_LEAF, PH_LEAF_WAIT, PH_COMBINE,
  PH_ITER_END, PH_HMC_STEP, PH_HMC_WAIT, PH_HMC_END
};

__device__ __forceinline__ Key K(const uint32_t* k) { return Key{k[0], k[1]}; }
__device__ __forceinline__ void setK(uint32_t* d, Key k) { d[0] = k.k0; d[1] = k.k1; }

struct ChainCtx {
  float* vb;       // chain's vector block [nvec][Dp]
  double* wel;     // Welford [2][Dp] (mean, m2) fp64
  int lane, D, Dp;
  __device__ float* v(int slot) const { return vb + (size_t)slot * Dp; }
};

#define FOR_D(d) for (int d = cx.lane; d < cx.D; d += NAX_WAVE)

__device__ __forceinline__ void vcopy(const ChainCtx& cx, int dst, int src) {
  float* a = cx.v(dst);
  const float* b = cx.v(src);
  FOR_D(d) a[d] = b[d];
}

__device__ double kinetic(const ChainCtx& cx, int r_slot) {
  const float* r = cx.v(r_slot);
  const float* m = cx.v(V_IMM);
  double acc = 0.0;
  FOR_D(d) { double rv = r[d]; acc += (double)m[d] * rv * rv; }
  return 0.5 * wave_sum(acc);
}

// _is_turning (hmc_util.py:710-746) with r_sum given as rs - rs_ck + r_ck when
// rs_ck != nullptr (the iterative subtree form, hmc_util.py:961-981).
__device__ bool is_turning(const ChainCtx& cx, const float* rl, const float* rr,
                           const float* rs, const float* rs_ck, const float* r_ck) {
  const float* m = cx.v(V_IMM);
  double a = 0.0, b = 0.0;
  FOR_D(d) {
    float s = rs[d];
    if (rs_ck) s = s - rs_ck[d] + r_ck[d];
    float rlv = rl[d], rrv = rr[d];
    float c = s - (rlv + rrv) * 0.5f;
    a += (double)(m[d] * rlv) * c;
    b += (double)(m[d] * rrv) * c;
  }
  a = wave_sum(a);
  b = wave_sum(b);
  return (a <= 0.0) || (b <= 0.0);
}

// momentum_generator (hmc.py:92-110), diagonal mass: sqrt(M) * normal(key, (D,))
__device__ void draw_momentum(const ChainCtx& cx, int dst, Key k, bool part) {
  float* r = cx.v(dst);
  const float* s = cx.v(V_MSQRT);
  FOR_D(d) r[d] = s[d] * normal_vec(k, (uint32_t)d, (uint32_t)cx.D, part);
}

// First half of velocity_verlet (hmc_util.py:289-297) on an edge, in place:
// r <- r - eps/2 g; z <- z + eps M^-1 r; z is also written to the request.
__device__ void leapfrog_half(const ChainCtx& cx, int zs, int rs, int gs, float eps,
                              float* zreq) {
  float* z = cx.v(zs);
  float* r = cx.v(rs);
  const float* g = cx.v(gs);
  const float* m = cx.v(V_IMM);
  const float h = 0.5f * eps;
  FOR_D(d) {
    float rh = r[d] - h * g[d];
    float zn = z[d] + eps * (m[d] * rh);
    r[d] = rh;
    z[d] = zn;
    zreq[d] = zn;
  }
}

// Second half (hmc_util.py:298-309): g <- grad(z); r <- r - eps/2 g.
__device__ void leapfrog_finish(const ChainCtx& cx, int rs, int gs, float eps,
                                const float* gin) {
  float* r = cx.v(rs);
  float* g = cx.v(gs);
  const float h = 0.5f * eps;
  FOR_D(d) {
    float gv = gin[d];
    g[d] = gv;
    r[d] = r[d] - h * gv;
  }
}

// ---- warmup adaptation (hmc_util.py:60-130, 133-239, 518-707) ---------------

__device__ void da_init(ChainState& S, double prox) {
  S.da_x = 0.0; S.da_xavg = 0.0; S.da_gavg = 0.0; S.da_t = 0; S.da_prox = prox;
}

__device__ void welford_reset(const ChainCtx& cx, ChainState& S) {
  double* mean = cx.wel;
  double* m2 = cx.wel + cx.Dp;
  FOR_D(d) { mean[d] = 0.0; m2[d] = 0.0; }
  S.welford_n = 0;
}

__device__ void adapt_update(const EngineCfg& cfg, const ChainCtx& cx, ChainState& S, int t,
                             double accept_prob) {
  const bool part = cfg.partitionable != 0;
  Key ak = K(S.adapt_key);
  setK(S.adapt_key, split_key(ak, 0, 2, part));  // rng_key_ss = split(...)[1] is unused
  if (cfg.adapt_step_size) {
    const double t0 = 10.0, kappa = 0.75, gamma = 0.05;
    double g = cfg.target_accept - accept_prob;
    int tt = S.da_t + 1;
    double gavg = (1.0 - 1.0 / (tt + t0)) * S.da_gavg + g / (tt + t0);
    double x = S.da_prox - sqrt((double)tt) / gamma * gavg;
    double w = pow((double)tt, -kappa);
    double xavg = (1.0 - w) * S.da_xavg + w * x;
    S.da_t = tt; S.da_gavg = gavg; S.da_x = x; S.da_xavg = xavg;
    double ss = (t == cfg.num_warmup - 1) ? exp(xavg) : exp(x);
    ss = fmin(fmax(ss, 2.2250738585072014e-308), 1.7976931348623157e308);
    S.step_size = ss;
  }
  const int widx = S.window_idx;
  const bool middle = (0 < widx) && (widx < cfg.num_windows - 1);
  if (cfg.adapt_mass_matrix && middle) {
    double* mean = cx.wel;
    double* m2 = cx.wel + cx.Dp;
    const float* z = cx.v(V_Z);
    const int n = S.welford_n + 1;
    FOR_D(d) {
      double x = z[d];
      double dpre = x - mean[d];
      double mn = mean[d] + dpre / n;
      m2[d] += dpre * (x - mn);
      mean[d] = mn;
    }
    S.welford_n = n;
  }
  const bool at_end = (t == cfg.window_end[widx]);
  if (at_end) S.window_idx = widx + 1;
  if (at_end && middle) {
    if (cfg.adapt_mass_matrix) {
      double* mean = cx.wel;
      double* m2 = cx.wel + cx.Dp;
      float* imm = cx.v(V_IMM);
      float* msq = cx.v(V_MSQRT);
      const double n = (double)S.welford_n;
      FOR_D(d) {
        double cov = m2[d] / (n - 1.0);
        if (cfg.regularize) cov = (n / (n + 5.0)) * cov + 1e-3 * (5.0 / (n + 5.0));
        imm[d] = (float)cov;
        msq[d] = (float)(1.0 / sqrt(cov));
        mean[d] = 0.0;
        m2[d] = 0.0;
      }
      S.welford_n = 0;
    }
    if (cfg.adapt_step_size) da_init(S, log(10.0) + log(S.step_size));
  }
}

// ---- collection (fori_collect, numpyro/util.py:384-403) ---------------------
__device__ void collect(const EngineCfg& cfg, const Collect& col, const ChainCtx& cx,
                        const ChainState& S, int c, int64_t it) {
  if (col.size <= 0) return;
  int64_t rel = it - col.start_idx;
  if (rel < 0) return;
  if ((rel % col.thinning) != col.thinning - 1) return;
  int64_t slot = rel / col.thinning;
  if (slot >= col.size) return;
  const size_t cs = (size_t)c * col.size + slot;
  if (col.z) {
    float* dst = col.z + cs * (size_t)cx.D;
    const float* z = cx.v(V_Z);
    FOR_D(d) dst[d] = z[d];
  }
  if (cx.lane == 0) {
    if (col.potential_energy) col.potential_energy[cs] = S.pe;
    if (col.energy) col.energy[cs] = S.energy;
    if (col.accept_prob) col.accept_prob[cs] = (float)S.accept_prob;
    if (col.mean_accept_prob) col.mean_accept_prob[cs] = (float)S.mean_accept_prob;
    if (col.num_steps) col.num_steps[cs] = S.num_steps;
    if (col.diverging) col.diverging[cs] = (uint8_t)S.diverging;
    if (col.step_size) col.step_size[cs] = (float)S.step_size;
  }
}

__device__ __forceinline__ double nan_to_inf(double x) { return isnan(x) ? INFINITY : x; }
__device__ __forceinline__ double clip_exp_neg(double delta) {
  // clip(exp(-delta), None, 1)
  return delta <= 0.0 ? 1.0 : exp(-delta);
}

// ---- the state machine -------------------------------------------------------
__global__ void __launch_bounds__(256)
advance_kernel(EngineCfg cfg, Collect col, int64_t target, ChainState* __restrict__ states,
               float* __restrict__ vec, double* __restrict__ welf, float* __restrict__ zreq,
               const float* __restrict__ gout, const double* __restrict__ peout,
               int* __restrict__ pending, int64_t* __restrict__ grad_evals) {
  const int c = blockIdx.x * (blockDim.x / NAX_WAVE) + threadIdx.x / NAX_WAVE;
  if (c >= cfg.C) return;
  ChainCtx cx;
  cx.lane = threadIdx.x & (NAX_WAVE - 1);
  cx.D = cfg.D;
  cx.Dp = cfg.Dp;
  cx.vb = vec + (size_t)c * cfg.nvec * cfg.Dp;
  cx.wel = welf + (size_t)c * 2 * cfg.Dp;
  const bool part = cfg.partitionable != 0;
  float* zr = zreq + (size_t)c * cfg.Dp;
  const float* gi = gout + (size_t)c * cfg.Dp;
  const int CKR = V_CK, CKS = V_CK + cfg.max_depth_alloc;

  ChainState S = states[c];
  bool need_grad = false;
  int evals = 0;

  while (!need_grad) {
    switch (S.phase) {
      case PH_INIT_WAIT: {  // init_kernel (hmc.py:338-349) after value_and_grad(z0)
        float* g = cx.v(V_G);
        FOR_D(d) g[d] = gi[d];
        S.pe = peout[c];
        S.energy = S.pe + kinetic(cx, V_R0);
        S.phase = PH_ITER_START;
        break;
      }
      case PH_ITER_START: {  // sample_kernel (hmc.py:459-490)
        if ((int64_t)S.i >= target) goto park;
        Key k0 = K(S.rng_key);
        Key kn = split_key(k0, 0, 3, part);
        Key kmom = split_key(k0, 1, 3, part);
        Key ktr = split_key(k0, 2, 3, part);
        setK(S.rng_key, kn);
        const bool warm = S.i < cfg.num_warmup;
        if (cfg.algo == 0) {
          draw_momentum(cx, V_RL, kmom, part);
          // tree = TreeInfo(z, r, z_grad, z, r, z_grad, z, pe, z_grad, energy, ...)
          {
            const float* z = cx.v(V_Z);
            const float* g = cx.v(V_G);
            float* zl = cx.v(V_ZL); float* gl = cx.v(V_GL);
            float* zrr = cx.v(V_ZR); float* rrr = cx.v(V_RR); float* grr = cx.v(V_GR);
            float* zp = cx.v(V_ZP); float* gp = cx.v(V_GP); float* rs = cx.v(V_RSUM);
            const float* rl = cx.v(V_RL);
            FOR_D(d) {
              float zv = z[d], gv = g[d], rv = rl[d];
              zl[d] = zv; gl[d] = gv; zrr[d] = zv; rrr[d] = rv; grr[d] = gv;
              zp[d] = zv; gp[d] = gv; rs[d] = rv;
            }
          }
          S.energy_current = S.pe + kinetic(cx, V_RL);
          S.peP = S.pe;
          S.energyP = S.energy_current;
          S.depth = 0; S.weight = 0.0; S.turning = 0; S.t_diverging = 0;
          S.sum_accept = 0.0; S.num_proposals = 0;
          S.max_depth_cur = warm ? cfg.max_depth_warmup : cfg.max_depth_post;
          setK(S.tree_key, ktr);
          S.phase = PH_DOUBLING;
        } else {
          // _hmc_next (hmc.py:364-414): integrate edge R from (z, r, z_grad)
          draw_momentum(cx, V_RR, kmom, part);
          vcopy(cx, V_ZR, V_Z);
          vcopy(cx, V_GR, V_G);
          S.energy_current = S.pe + kinetic(cx, V_RR);
          S.pe_edge = S.pe;
          int ns;
          if (cfg.fixed_num_steps > 0) ns = cfg.fixed_num_steps;
          else ns = (int)ceil(cfg.trajectory_length / S.step_size);
          S.hmc_num_steps = ns;
          S.hmc_eps = (cfg.trajectory_length > 0) ? cfg.trajectory_length / ns : S.step_size;
          S.hmc_step = 0;
          setK(S.trans_key, ktr);
          S.phase = PH_HMC_STEP;
        }
        break;
      }
      case PH_DOUBLING: {  // build_tree's while_loop (hmc_util.py:1156-1179)
        if (S.depth < S.max_depth_cur && !S.turning && !S.t_diverging) {
          Key tk = K(S.tree_key);
          Key kn = split_key(tk, 0, 3, part);
          Key kdir = split_key(tk, 1, 3, part);
          Key kdbl = split_key(tk, 2, 3, part);
          setK(S.tree_key, kn);
          S.going_right = bernoulli(kdir, 0.5, part) ? 1 : 0;
          setK(S.sub_key, split_key(kdbl, 0, 2, part));  // _double_tree (hmc_util.py:920)
          setK(S.top_key, split_key(kdbl, 1, 2, part));
          S.s_num = 0;
          S.s_turning = 0;
          S.s_diverging = 0;
          S.phase = PH_LEAF;
        } else {
          S.phase = PH_ITER_END;
        }
        break;
      }
      case PH_LEAF: {  // _iterative_build_subtree body (hmc_util.py:1003-1017)
        Key sk = K(S.sub_key);
        setK(S.leaf_key, split_key(sk, 1, 2, part));
        setK(S.sub_key, split_key(sk, 0, 2, part));
        const float eps = (float)(S.going_right ? S.step_size : -S.step_size);
        if (S.going_right) leapfrog_half(cx, V_ZR, V_RR, V_GR, eps, zr);
        else leapfrog_half(cx, V_ZL, V_RL, V_GL, eps, zr);
        S.phase = PH_LEAF_WAIT;
        need_grad = true;
        break;
      }
      case PH_LEAF_WAIT: {  // _build_basetree (hmc_util.py:851-894) + combine + ckpts
        evals++;
        const int zs = S.going_right ? V_ZR : V_ZL;
        const int rs = S.going_right ? V_RR : V_RL;
        const int gs = S.going_right ? V_GR : V_GL;
        const float eps = (float)(S.going_right ? S.step_size : -S.step_size);
        leapfrog_finish(cx, rs, gs, eps, gi);
        const double pe_new = peout[c];
        const double e_new = pe_new + kinetic(cx, rs);
        const double delta = nan_to_inf(e_new - S.energy_current);
        const double leaf_w = -delta;
        const int leaf_div = delta > cfg.max_delta_energy;
        const double acc = clip_exp_neg(delta);
        const int leaf_idx = S.s_num;
        const float* znew = cx.v(zs);
        const float* rnew = cx.v(rs);
        const float* gnew = cx.v(gs);
        float* szp = cx.v(V_SZP);
        float* sgp = cx.v(V_SGP);
        float* srs = cx.v(V_SRSUM);
        if (leaf_idx == 0) {
          FOR_D(d) { szp[d] = znew[d]; sgp[d] = gnew[d]; srs[d] = rnew[d]; }
          S.s_peP = pe_new; S.s_energyP = e_new; S.s_weight = leaf_w;
          S.s_diverging = leaf_div; S.s_sum_accept = acc; S.s_num = 1;
        } else {  // _combine_tree(..., biased_transition=False) (hmc_util.py:767-848)
          FOR_D(d) srs[d] += rnew[d];
          const double p = expit(leaf_w - S.s_weight);
          if (bernoulli(K(S.leaf_key), p, part)) {
            FOR_D(d) { szp[d] = znew[d]; sgp[d] = gnew[d]; }
            S.s_peP = pe_new; S.s_energyP = e_new;
          }
          S.s_weight = log_add_exp(S.s_weight, leaf_w);
          S.s_diverging = leaf_div;
          S.s_sum_accept += acc;
          S.s_num += 1;
        }
        // _leaf_idx_to_ckpt_idxs (hmc_util.py:941-958)
        const int idx_max = __popc((unsigned)leaf_idx >> 1);
        const int nsub = __popc((unsigned)((~leaf_idx & (leaf_idx + 1)) - 1));
        const int idx_min = idx_max - nsub + 1;
        if ((leaf_idx & 1) == 0) {
          float* ckr = cx.v(CKR + idx_max);
          float* cks = cx.v(CKS + idx_max);
          FOR_D(d) { ckr[d] = rnew[d]; cks[d] = srs[d]; }
        }
        bool turning = false;
        for (int i = idx_max; i >= idx_min && !turning; --i)
          turning = is_turning(cx, cx.v(CKR + i), rnew, srs, cx.v(CKS + i), cx.v(CKR + i));
        S.s_turning = turning ? 1 : 0;
        if (S.s_num < (1 << S.depth) && !turning && !S.s_diverging) S.phase = PH_LEAF;
        else S.phase = PH_COMBINE;
        break;
      }
      case PH_COMBINE: {  // _double_tree's _combine_tree(..., True) (hmc_util.py:936-938)
        float* rsum = cx.v(V_RSUM);
        const float* srs = cx.v(V_SRSUM);
        FOR_D(d) rsum[d] += srs[d];
        double p = 0.0;
        if (!S.s_turning && !S.s_diverging) p = fmin(exp(fmin(S.s_weight - S.weight, 700.0)), 1.0);
        const bool turn = S.s_turning || is_turning(cx, cx.v(V_RL), cx.v(V_RR), rsum, nullptr, nullptr);
        if (bernoulli(K(S.top_key), p, part)) {
          vcopy(cx, V_ZP, V_SZP);
          vcopy(cx, V_GP, V_SGP);
          S.peP = S.s_peP;
          S.energyP = S.s_energyP;
        }
        S.depth += 1;
        S.weight = log_add_exp(S.weight, S.s_weight);
        S.turning = turn ? 1 : 0;
        S.t_diverging = S.s_diverging;
        S.sum_accept += S.s_sum_accept;
        S.num_proposals += S.s_num;
        S.phase = PH_DOUBLING;
        break;
      }
      case PH_HMC_STEP: {
        if (S.hmc_step < S.hmc_num_steps) {
          leapfrog_half(cx, V_ZR, V_RR, V_GR, (float)S.hmc_eps, zr);
          S.phase = PH_HMC_WAIT;
          need_grad = true;
        } else {
          S.phase = PH_HMC_END;
        }
        break;
      }
      case PH_HMC_WAIT: {
        evals++;
        leapfrog_finish(cx, V_RR, V_GR, (float)S.hmc_eps, gi);
        S.pe_edge = peout[c];
        S.hmc_step += 1;
        S.phase = PH_HMC_STEP;
        break;
      }
      case PH_HMC_END: {  // hmc.py:398-414
        const double e_new = S.pe_edge + kinetic(cx, V_RR);
        const double delta = nan_to_inf(e_new - S.energy_current);
        const double acc = clip_exp_neg(delta);
        S.t_diverging = delta > cfg.max_delta_energy;
        S.sum_accept = acc;
        S.num_proposals = S.hmc_num_steps;
        if (bernoulli(K(S.trans_key), acc, part)) {
          vcopy(cx, V_ZP, V_ZR);
          vcopy(cx, V_GP, V_GR);
          S.peP = S.pe_edge;
          S.energyP = e_new;
        } else {
          vcopy(cx, V_ZP, V_Z);
          vcopy(cx, V_GP, V_G);
          S.peP = S.pe;
          S.energyP = S.energy_current;
        }
        S.phase = PH_ITER_END;
        break;
      }
      case PH_ITER_END: {  // _nuts_next (hmc.py:441-455) + sample_kernel (hmc.py:492-530)
        const double ap = (cfg.algo == 0) ? S.sum_accept / S.num_proposals : S.sum_accept;
        S.accept_prob = ap;
        S.num_steps = S.num_proposals;
        S.diverging = S.t_diverging;
        vcopy(cx, V_Z, V_ZP);
        vcopy(cx, V_G, V_GP);
        S.pe = S.peP;
        S.energy = S.energyP;
        const int it = S.i;
        if (it < cfg.num_warmup) adapt_update(cfg, cx, S, it, ap);
        const int itr = it + 1;
        const int n = (it < cfg.num_warmup) ? itr : itr - cfg.num_warmup;
        S.mean_accept_prob = S.mean_accept_prob + (ap - S.mean_accept_prob) / n;
        S.i = itr;
        collect(cfg, col, cx, S, c, it);
        S.phase = PH_ITER_START;
        break;
      }
      default:
        goto park;
    }
  }
park:
  if (cx.lane == 0) {
    states[c] = S;
    if (need_grad) atomicAdd(pending, 1);
    if (evals && grad_evals) atomicAdd((unsigned long long*)grad_evals, (unsigned long long)evals);
  }
}

// init_kernel (hmc.py:193-362) + warmup_adapter.init_fn (hmc_util.py:552-586)
__global__ void __launch_bounds__(256)
init_kernel(EngineCfg cfg, ChainState* __restrict__ states, float* __restrict__ vec,
            double* __restrict__ welf, float* __restrict__ zreq, const float* __restrict__ z0,
            const uint32_t* __restrict__ keys, const float* __restrict__ imm0) {
  const int c = blockIdx.x * (blockDim.x / NAX_WAVE) + threadIdx.x / NAX_WAVE;
  if (c >= cfg.C) return;
  ChainCtx cx;
  cx.lane = threadIdx.x & (NAX_WAVE - 1);
  cx.D = cfg.D;
  cx.Dp = cfg.Dp;
  cx.vb = vec + (size_t)c * cfg.nvec * cfg.Dp;
  cx.wel = welf + (size_t)c * 2 * cfg.Dp;
  const bool part = cfg.partitionable != 0;
  ChainState S;
  memset(&S, 0, sizeof(S));
  Key k = Key{keys[2 * c], keys[2 * c + 1]};
  Key k_hmc = split_key(k, 0, 3, part);
  Key k_wa = split_key(k, 1, 3, part);
  Key k_mom = split_key(k, 2, 3, part);
  setK(S.adapt_key, split_key(k_wa, 0, 2, part));
  setK(S.rng_key, k_hmc);
  float* imm = cx.v(V_IMM);
  float* msq = cx.v(V_MSQRT);
  FOR_D(d) {
    float m = imm0 ? imm0[d] : 1.0f;
    imm[d] = m;
    msq[d] = imm0 ? (float)(1.0 / sqrt((double)m)) : 1.0f;
  }
  S.step_size = cfg.trajectory_length;  // overwritten by the host value below
  welford_reset(cx, S);
  draw_momentum(cx, V_R0, k_mom, part);
  float* z = cx.v(V_Z);
  float* zr = zreq + (size_t)c * cfg.Dp;
  const float* zi = z0 + (size_t)c * cfg.D;
  FOR_D(d) { z[d] = zi[d]; zr[d] = zi[d]; }
  S.phase = PH_INIT_WAIT;
  S.i = 0;
  if (cx.lane == 0) states[c] = S;
}

__global__ void set_step_size_kernel(ChainState* states, int C, double step) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C) {
    states[c].step_size = step;
    states[c].da_prox = log(10.0 * step);
    states[c].da_x = 0.0; states[c].da_xavg = 0.0; states[c].da_gavg = 0.0; states[c].da_t = 0;
  }
}

__global__ void zero_ints_kernel(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---- state import / export -----------------------------------------------------
__global__ void state_vec_io(int C, int D, int Dp, int nvec, float* vec, int slot,
                             float* buf, int to_vec) {
  int c = blockIdx.y;
  int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C || d >= D) return;
  float* v = vec + ((size_t)c * nvec + slot) * Dp + d;
  if (to_vec) *v = buf[(size_t)c * D + d];
  else buf[(size_t)c * D + d] = *v;
}

// ---- standalone leapfrog (parity entry point) ------------------------------------
__global__ void lf_half_kernel(int C, int D, int ldz, float* z, float* r, const float* g,
                               const float* imm, float eps, float* zreq, int ldq) {
  int c = blockIdx.y;
  int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C || d >= D) return;
  size_t i = (size_t)c * ldz + d;
  float rh = r[i] - 0.5f * eps * g[i];
  float zn = z[i] + eps * (imm[d] * rh);
  r[i] = rh;
  z[i] = zn;
  zreq[(size_t)c * ldq + d] = zn;
}
__global__ void lf_finish_kernel(int C, int D, int ldz, float* r, float* g, const float* gin,
                                 int ldq, float eps, double* pe, const double* pein) {
  int c = blockIdx.y;
  int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C || d >= D) return;
  size_t i = (size_t)c * ldz + d;
  float gv = gin[(size_t)c * ldq + d];
  g[i] = gv;
  r[i] = r[i] - 0.5f * eps * gv;
  if (d == 0) pe[c] = pein[c];
}

// ---- random test kernels -----------------------------------------------------------
__global__ void tf_kernel(const uint32_t* key, const uint32_t* x0, const uint32_t* x1,
                          uint32_t* o0, uint32_t* o1, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t a = x0[i], b = x1[i];
  threefry2x32(Key{key[0], key[1]}, a, b);
  o0[i] = a;
  o1[i] = b;
}
__global__ void split_kernel(const uint32_t* key, int num, int part, uint32_t* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num) return;
  Key k = split_key(Key{key[0], key[1]}, (uint32_t)i, (uint32_t)num, part != 0);
  out[2 * i] = k.k0;
  out[2 * i + 1] = k.k1;
}
// MCMC.run: split(key, C) (mcmc.py:670-671), then HMC.init: split(k_c)[0]
// (hmc.py:744-750) -> the rng_key handed to init_kernel for chain c.
__global__ void chain_keys_kernel(const uint32_t* key, int C, int part, uint32_t* out) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  Key k = Key{key[0], key[1]};
  Key kc = (C > 1) ? split_key(k, (uint32_t)c, (uint32_t)C, part != 0) : k;
  Key ki = split_key(kc, 0, 2, part != 0);
  out[2 * c] = ki.k0;
  out[2 * c + 1] = ki.k1;
}
__global__ void normal_kernel(const uint32_t* key, int n, int part, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = normal_vec(Key{key[0], key[1]}, (uint32_t)i, (uint32_t)n, part != 0);
}

// ---- host side ---------------------------------------------------------------------
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

}  // namespace nax

using namespace nax;

struct nax_sampler {
  EngineCfg cfg;
  nax_config user;
  hipStream_t stream = nullptr;
  Potential* pot = nullptr;
  bool external = false;
  bool initialized = false;
  float* vec = nullptr;
  double* wel = nullptr;
  ChainState* st = nullptr;
  float* zreq = nullptr;
  float* gout = nullptr;
  double* peout = nullptr;
  int* ctr = nullptr;          // device counters [kBatch]
  int64_t* gevals = nullptr;   // device
  int* ctr_host = nullptr;     // pinned
  int64_t* gevals_host = nullptr;
  void (*progress_cb)(void*, int64_t, int32_t) = nullptr;
  void* progress_user = nullptr;
  int64_t progress_every = 0;
  bool timing = false;
  double timing_ms = 0.0;
  int64_t timing_n = 0;
  std::vector<hipEvent_t> ev;
};

static const int kBatch = 16;

#define NAX_HIP(x)                                                                 \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      set_error(std::string(#x) + ": " + hipGetErrorString(e_));                   \
      return NAX_EHIP;                                                             \
    }                                                                              \
  } while (0)

extern "C" const char* nax_last_error(void) { return g_err.c_str(); }

static std::vector<int> schedule_ends(int num_steps) {
  // build_adaptation_schedule (numpyro/infer/hmc_util.py:387-436)
  std::vector<int> ends;
  if (num_steps <= 0) return ends;
  if (num_steps < 20) {
    ends.push_back(num_steps - 1);
    return ends;
  }
  int sb = 75, eb = 50, iw = 25;
  if (sb + eb + iw > num_steps) {
    sb = (int)(0.15 * num_steps);
    eb = (int)(0.1 * num_steps);
    iw = num_steps - sb - eb;
  }
  ends.push_back(sb - 1);
  int end_start = num_steps - eb;
  int next_size = iw, next_start = sb;
  while (next_start < end_start) {
    int cs = next_start, csz = next_size;
    if (3 * csz <= end_start - cs) next_size = 2 * csz;
    else csz = end_start - cs;
    next_start = cs + csz;
    ends.push_back(next_start - 1);
  }
  ends.push_back(num_steps - 1);
  return ends;
}

extern "C" int nax_create(const nax_config* c, nax_sampler** out) {
  if (!c || !out) { set_error("nax_create: null argument"); return NAX_EINVAL; }
  if (c->num_chains <= 0 || c->dim <= 0) { set_error("num_chains and dim must be positive"); return NAX_EINVAL; }
  if (c->algo != 0 && c->algo != 1) { set_error("`algo` must be one of `HMC` or `NUTS`."); return NAX_EINVAL; }
  if (c->max_tree_depth <= 0 || c->max_tree_depth_warmup <= 0 || c->max_tree_depth > 30 ||
      c->max_tree_depth_warmup > 30) {
    set_error("max_tree_depth must be in [1, 30]");
    return NAX_EINVAL;
  }
  if (c->algo == 1 && c->num_steps <= 0 && !(c->trajectory_length > 0)) {
    set_error("At least one of `num_steps` or `trajectory_length` must be specified.");
    return NAX_EINVAL;
  }
  if (!(c->step_size > 0)) { set_error("step_size must be positive"); return NAX_EINVAL; }
  std::unique_ptr<nax_sampler> h(new nax_sampler());
  h->user = *c;
  EngineCfg& e = h->cfg;
  memset(&e, 0, sizeof(e));
  e.C = c->num_chains;
  e.D = c->dim;
  e.Dp = (c->dim + 63) / 64 * 64;
  e.algo = c->algo;
  e.num_warmup = c->num_warmup;
  e.max_depth_warmup = c->max_tree_depth_warmup;
  e.max_depth_post = c->max_tree_depth;
  e.max_depth_alloc = std::max(c->max_tree_depth, c->max_tree_depth_warmup);
  e.nvec = V_CK + 2 * e.max_depth_alloc;
  e.adapt_step_size = c->adapt_step_size;
  e.adapt_mass_matrix = c->adapt_mass_matrix;
  e.regularize = c->regularize_mass_matrix;
  e.partitionable = c->partitionable_keys;
  e.fixed_num_steps = c->num_steps;
  e.trajectory_length = c->trajectory_length;
  e.target_accept = c->target_accept_prob;
  e.max_delta_energy = c->max_delta_energy > 0 ? c->max_delta_energy : 1000.0;
  std::vector<int> ends = schedule_ends(c->num_warmup);
  if ((int)ends.size() > kMaxWindows) { set_error("too many adaptation windows"); return NAX_EINVAL; }
  e.num_windows = (int)ends.size();
  for (size_t i = 0; i < ends.size(); ++i) e.window_end[i] = ends[i];
  for (int i = e.num_windows; i < kMaxWindows; ++i) e.window_end[i] = -1;
  if (e.num_windows == 0) { e.num_windows = 1; e.window_end[0] = -1; }

  NAX_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  const size_t C = e.C, Dp = e.Dp;
  NAX_HIP(hipMalloc(&h->vec, C * e.nvec * Dp * sizeof(float)));
  NAX_HIP(hipMalloc(&h->wel, C * 2 * Dp * sizeof(double)));
  NAX_HIP(hipMalloc(&h->st, C * sizeof(ChainState)));
  NAX_HIP(hipMalloc(&h->zreq, C * Dp * sizeof(float)));
  NAX_HIP(hipMalloc(&h->gout, C * Dp * sizeof(float)));
  NAX_HIP(hipMalloc(&h->peout, C * sizeof(double)));
  NAX_HIP(hipMalloc(&h->ctr, kBatch * sizeof(int)));
  NAX_HIP(hipMalloc(&h->gevals, sizeof(int64_t)));
  NAX_HIP(hipHostMalloc(&h->ctr_host, sizeof(int) * 2, hipHostMallocDefault));
  NAX_HIP(hipHostMalloc(&h->gevals_host, sizeof(int64_t), hipHostMallocDefault));
  NAX_HIP(hipMemsetAsync(h->vec, 0, C * e.nvec * Dp * sizeof(float), h->stream));
  NAX_HIP(hipMemsetAsync(h->zreq, 0, C * Dp * sizeof(float), h->stream));
  NAX_HIP(hipMemsetAsync(h->gout, 0, C * Dp * sizeof(float), h->stream));
  NAX_HIP(hipStreamSynchronize(h->stream));
  *out = h.release();
  return NAX_OK;
}

extern "C" void nax_destroy(nax_sampler* h) {
  if (!h) return;
  if (h->stream) hipStreamSynchronize(h->stream);
  delete h->pot;
  hipFree(h->vec); hipFree(h->wel); hipFree(h->st); hipFree(h->zreq); hipFree(h->gout);
  hipFree(h->peout); hipFree(h->ctr); hipFree(h->gevals);
  hipHostFree(h->ctr_host); hipHostFree(h->gevals_host);
  for (auto e : h->ev) hipEventDestroy(e);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

extern "C" int nax_model_logistic(nax_sampler* h, const float* x, const float* y, int64_t n,
                                  int32_t d, double prior_scale, int32_t x_dtype) {
  return nax_model_logistic_ex(h, x, y, n, d, prior_scale, x_dtype, 0);
}

extern "C" int nax_model_logistic_ex(nax_sampler* h, const float* x, const float* y, int64_t n,
                                     int32_t d, double prior_scale, int32_t x_dtype,
                                     int32_t kernel) {
  if (!h || !x || !y) { set_error("nax_model_logistic: null argument"); return NAX_EINVAL; }
  if (d != h->cfg.D) { set_error("logistic model dim does not match the sampler dim"); return NAX_EINVAL; }
  if (n <= 0) { set_error("logistic model needs n > 0 rows"); return NAX_EINVAL; }
  if (!(prior_scale > 0)) { set_error("prior_scale must be positive"); return NAX_EINVAL; }
  std::string err;
  if (kernel < 0 || kernel > 2) { set_error("kernel must be 0 (auto), 1 (tile) or 2 (matrix cores)"); return NAX_EINVAL; }
  Potential* p = make_logistic(x, y, n, d, prior_scale, x_dtype, kernel, h->cfg.C, h->stream, &err);
  if (!p) { set_error(err); return NAX_EINVAL; }
  delete h->pot;
  h->pot = p;
  h->external = false;
  return NAX_OK;
}

extern "C" int nax_model_external(nax_sampler* h) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  delete h->pot;
  h->pot = nullptr;
  h->external = true;
  return NAX_OK;
}

// Evaluates the potential for every chain's request.  With timing on, the
// main kernel of launch k of a batch is bracketed by events ev[2k], ev[2k+1];
// collect_timing() reads them after the batch has synchronised.
static int eval_potential(nax_sampler* h, int k = 0) {
  const EngineCfg& e = h->cfg;
  if (h->timing) {
    if ((int)h->ev.size() < 2 * kBatch) {
      for (auto x : h->ev) hipEventDestroy(x);
      h->ev.assign(2 * kBatch, nullptr);
      for (auto& x : h->ev) NAX_HIP(hipEventCreate(&x));
    }
    h->pot->ev_begin = h->ev[2 * k];
    h->pot->ev_end = h->ev[2 * k + 1];
  } else {
    h->pot->ev_begin = h->pot->ev_end = nullptr;
  }
  NAX_HIP(h->pot->eval(h->stream, h->zreq, e.Dp, e.C, h->peout, h->gout, e.Dp));
  h->pot->ev_begin = h->pot->ev_end = nullptr;
  return NAX_OK;
}

static int collect_timing(nax_sampler* h, int nlaunch) {
  if (!h->timing) return NAX_OK;
  for (int k = 0; k < nlaunch; ++k) {
    float ms = 0;
    NAX_HIP(hipEventElapsedTime(&ms, h->ev[2 * k], h->ev[2 * k + 1]));
    h->timing_ms += ms;
    h->timing_n += 1;
  }
  return NAX_OK;
}

static Collect to_collect(const nax_collect* c) {
  Collect col;
  memset(&col, 0, sizeof(col));
  if (!c) return col;
  col.start_idx = c->start_idx;
  col.thinning = c->thinning > 0 ? c->thinning : 1;
  col.size = c->size;
  col.z = c->d_z;
  col.potential_energy = c->d_potential_energy;
  col.energy = c->d_energy;
  col.accept_prob = c->d_accept_prob;
  col.mean_accept_prob = c->d_mean_accept_prob;
  col.num_steps = c->d_num_steps;
  col.diverging = c->d_diverging;
  col.step_size = c->d_step_size;
  return col;
}

static int launch_advance(nax_sampler* h, const Collect& col, int64_t target, int slot) {
  const EngineCfg& e = h->cfg;
  dim3 grid((e.C + 3) / 4);
  hipLaunchKernelGGL(advance_kernel, grid, dim3(256), 0, h->stream, e, col, target, h->st,
                     h->vec, h->wel, h->zreq, h->gout, h->peout, h->ctr + slot, h->gevals);
  NAX_HIP(hipGetLastError());
  return NAX_OK;
}

extern "C" int nax_init(nax_sampler* h, const float* z0, const uint32_t* keys,
                        const float* imm) {
  if (!h || !z0 || !keys) { set_error("nax_init: null argument"); return NAX_EINVAL; }
  if (!h->pot && !h->external) { set_error("nax_init: no model set"); return NAX_ESTATE; }
  const EngineCfg& e = h->cfg;
  uint32_t* dkeys = nullptr;
  NAX_HIP(hipMallocAsync((void**)&dkeys, sizeof(uint32_t) * 2 * e.C, h->stream));
  NAX_HIP(hipMemcpyAsync(dkeys, keys, sizeof(uint32_t) * 2 * e.C, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(init_kernel, dim3((e.C + 3) / 4), dim3(256), 0, h->stream, e, h->st, h->vec,
                     h->wel, h->zreq, z0, dkeys, imm);
  NAX_HIP(hipGetLastError());
  hipLaunchKernelGGL(set_step_size_kernel, dim3((e.C + 255) / 256), dim3(256), 0, h->stream,
                     h->st, e.C, h->user.step_size);
  NAX_HIP(hipGetLastError());
  NAX_HIP(hipFreeAsync(dkeys, h->stream));
  h->initialized = true;
  if (!h->external) {
    // value_and_grad at z0, then complete the init state (advance to target 0).
    int rc = eval_potential(h);
    if (rc) return rc;
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, h->stream, h->ctr, kBatch);
    Collect col;
    memset(&col, 0, sizeof(col));
    rc = launch_advance(h, col, 0, 0);
    if (rc) return rc;
    NAX_HIP(hipStreamSynchronize(h->stream));
  }
  return NAX_OK;
}

extern "C" int nax_run(nax_sampler* h, int64_t target, const nax_collect* c, int64_t* gevals) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  if (!h->initialized) { set_error("nax_run before nax_init"); return NAX_ESTATE; }
  if (h->external || !h->pot) { set_error("nax_run needs a built-in model; use nax_advance"); return NAX_ESTATE; }
  Collect col = to_collect(c);
  NAX_HIP(hipMemsetAsync(h->gevals, 0, sizeof(int64_t), h->stream));
  int64_t ticks = 0;
  for (;;) {
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, h->stream, h->ctr, kBatch);
    for (int k = 0; k < kBatch; ++k) {
      int rc = launch_advance(h, col, target, k);
      if (rc) return rc;
      rc = eval_potential(h, k);
      if (rc) return rc;
    }
    NAX_HIP(hipMemcpyAsync(h->ctr_host, h->ctr + kBatch - 1, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    NAX_HIP(hipStreamSynchronize(h->stream));
    int rc = collect_timing(h, kBatch);
    if (rc) return rc;
    ticks += kBatch;
    if (h->progress_cb && h->progress_every > 0 && ticks % h->progress_every == 0)
      h->progress_cb(h->progress_user, ticks, h->ctr_host[0]);
    if (h->ctr_host[0] == 0) break;
  }
  if (gevals) {
    NAX_HIP(hipMemcpyAsync(h->gevals_host, h->gevals, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    NAX_HIP(hipStreamSynchronize(h->stream));
    *gevals = h->gevals_host[0];
  }
  return NAX_OK;
}

extern "C" int nax_advance(nax_sampler* h, int64_t target, const nax_collect* c, int32_t* pending) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  if (!h->initialized) { set_error("nax_advance before nax_init"); return NAX_ESTATE; }
  Collect col = to_collect(c);
  hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, h->stream, h->ctr, 1);
  int rc = launch_advance(h, col, target, 0);
  if (rc) return rc;
  NAX_HIP(hipMemcpyAsync(h->ctr_host, h->ctr, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  NAX_HIP(hipStreamSynchronize(h->stream));
  if (pending) *pending = h->ctr_host[0];
  return NAX_OK;
}

extern "C" int nax_buffers(nax_sampler* h, float** z, double** pe, float** g, int32_t* ld) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  if (z) *z = h->zreq;
  if (pe) *pe = h->peout;
  if (g) *g = h->gout;
  if (ld) *ld = h->cfg.Dp;
  return NAX_OK;
}

static int state_io(nax_sampler* h, const nax_state_view* v, bool to_dev) {
  const EngineCfg& e = h->cfg;
  dim3 grid((e.D + 127) / 128, e.C);
  struct { float* p; int slot; } vs[] = {{v->d_z, V_Z}, {v->d_z_grad, V_G},
                                         {v->d_inverse_mass_matrix, V_IMM},
                                         {v->d_mass_matrix_sqrt, V_MSQRT}};
  for (auto& x : vs) {
    if (!x.p) continue;
    hipLaunchKernelGGL(state_vec_io, grid, dim3(128), 0, h->stream, e.C, e.D, e.Dp, e.nvec,
                       h->vec, x.slot, x.p, to_dev ? 1 : 0);
    NAX_HIP(hipGetLastError());
  }
  std::vector<ChainState> hs(e.C);
  NAX_HIP(hipMemcpyAsync(hs.data(), h->st, sizeof(ChainState) * e.C, hipMemcpyDeviceToHost, h->stream));
  NAX_HIP(hipStreamSynchronize(h->stream));
  for (int c = 0; c < e.C; ++c) {
    ChainState& s = hs[c];
    if (to_dev) {
      if (v->i) s.i = v->i[c];
      if (v->potential_energy) s.pe = v->potential_energy[c];
      if (v->energy) s.energy = v->energy[c];
      if (v->accept_prob) s.accept_prob = v->accept_prob[c];
      if (v->mean_accept_prob) s.mean_accept_prob = v->mean_accept_prob[c];
      if (v->num_steps) s.num_steps = v->num_steps[c];
      if (v->diverging) s.diverging = v->diverging[c];
      if (v->step_size) s.step_size = v->step_size[c];
      if (v->rng_key) { s.rng_key[0] = v->rng_key[2 * c]; s.rng_key[1] = v->rng_key[2 * c + 1]; }
      if (s.phase != PH_INIT_WAIT) s.phase = PH_ITER_START;
    } else {
      if (v->i) v->i[c] = s.i;
      if (v->potential_energy) v->potential_energy[c] = s.pe;
      if (v->energy) v->energy[c] = s.energy;
      if (v->accept_prob) v->accept_prob[c] = s.accept_prob;
      if (v->mean_accept_prob) v->mean_accept_prob[c] = s.mean_accept_prob;
      if (v->num_steps) v->num_steps[c] = s.num_steps;
      if (v->diverging) v->diverging[c] = s.diverging;
      if (v->step_size) v->step_size[c] = s.step_size;
      if (v->rng_key) { v->rng_key[2 * c] = s.rng_key[0]; v->rng_key[2 * c + 1] = s.rng_key[1]; }
    }
  }
  if (to_dev) {
    NAX_HIP(hipMemcpyAsync(h->st, hs.data(), sizeof(ChainState) * e.C, hipMemcpyHostToDevice, h->stream));
    NAX_HIP(hipStreamSynchronize(h->stream));
  }
  return NAX_OK;
}

extern "C" int nax_get_state(nax_sampler* h, const nax_state_view* v) {
  if (!h || !v) { set_error("null argument"); return NAX_EINVAL; }
  if (!h->initialized) { set_error("no state before nax_init"); return NAX_ESTATE; }
  return state_io(h, v, false);
}

extern "C" int nax_set_state(nax_sampler* h, const nax_state_view* v) {
  if (!h || !v) { set_error("null argument"); return NAX_EINVAL; }
  if (!h->initialized) { set_error("nax_set_state before nax_init"); return NAX_ESTATE; }
  return state_io(h, v, true);
}

extern "C" int nax_potential_grad(nax_sampler* h, const float* z, int32_t ld, int32_t count,
                                  double* pe, float* g) {
  if (!h || !z || !pe || !g) { set_error("null argument"); return NAX_EINVAL; }
  if (!h->pot) { set_error("no built-in model set"); return NAX_ESTATE; }
  if (count <= 0 || ld < h->cfg.D) { set_error("bad count / ld"); return NAX_EINVAL; }
  NAX_HIP(h->pot->eval(h->stream, z, ld, count, pe, g, ld));
  NAX_HIP(hipStreamSynchronize(h->stream));
  return NAX_OK;
}

extern "C" int nax_leapfrog(nax_sampler* h, float* z, float* r, float* g, double* pe,
                            double step, int32_t L, const float* imm) {
  if (!h || !z || !r || !g || !pe || !imm) { set_error("null argument"); return NAX_EINVAL; }
  if (!h->pot) { set_error("no built-in model set"); return NAX_ESTATE; }
  const EngineCfg& e = h->cfg;
  dim3 grid((e.D + 127) / 128, e.C);
  for (int s = 0; s < L; ++s) {
    hipLaunchKernelGGL(lf_half_kernel, grid, dim3(128), 0, h->stream, e.C, e.D, e.D, z, r, g,
                       imm, (float)step, h->zreq, e.Dp);
    NAX_HIP(hipGetLastError());
    NAX_HIP(h->pot->eval(h->stream, h->zreq, e.Dp, e.C, h->peout, h->gout, e.Dp));
    hipLaunchKernelGGL(lf_finish_kernel, grid, dim3(128), 0, h->stream, e.C, e.D, e.D, r, g,
                       h->gout, e.Dp, (float)step, pe, h->peout);
    NAX_HIP(hipGetLastError());
  }
  NAX_HIP(hipStreamSynchronize(h->stream));
  return NAX_OK;
}

extern "C" int nax_kernel_timing(nax_sampler* h, int32_t enable, double* total_ms,
                                 int64_t* launches) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  if (total_ms) *total_ms = h->timing_ms;
  if (launches) *launches = h->timing_n;
  h->timing = enable != 0;
  h->timing_ms = 0.0;
  h->timing_n = 0;
  return NAX_OK;
}

extern "C" int nax_threefry2x32(const uint32_t* key, const uint32_t* x0, const uint32_t* x1,
                                uint32_t* o0, uint32_t* o1, int64_t n) {
  if (!key || !x0 || !x1 || !o0 || !o1 || n < 0) { set_error("bad argument"); return NAX_EINVAL; }
  if (n == 0) return NAX_OK;
  hipLaunchKernelGGL(tf_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, key, x0, x1, o0, o1, n);
  NAX_HIP(hipGetLastError());
  NAX_HIP(hipDeviceSynchronize());
  return NAX_OK;
}

extern "C" int nax_random_split(const uint32_t* key, int32_t num, int32_t part, uint32_t* out) {
  if (!key || !out || num <= 0) { set_error("bad argument"); return NAX_EINVAL; }
  hipLaunchKernelGGL(split_kernel, dim3((num + 255) / 256), dim3(256), 0, 0, key, num, part, out);
  NAX_HIP(hipGetLastError());
  NAX_HIP(hipDeviceSynchronize());
  return NAX_OK;
}

extern "C" int nax_random_normal(const uint32_t* key, int32_t n, int32_t part, float* out) {
  if (!key || !out || n <= 0) { set_error("bad argument"); return NAX_EINVAL; }
  hipLaunchKernelGGL(normal_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, key, n, part, out);
  NAX_HIP(hipGetLastError());
  NAX_HIP(hipDeviceSynchronize());
  return NAX_OK;
}

extern "C" int nax_chain_keys(const uint32_t* key, int32_t C, int32_t part, uint32_t* out) {
  if (!key || !out || C <= 0) { set_error("bad argument"); return NAX_EINVAL; }
  hipLaunchKernelGGL(chain_keys_kernel, dim3((C + 255) / 256), dim3(256), 0, 0, key, C, part, out);
  NAX_HIP(hipGetLastError());
  NAX_HIP(hipDeviceSynchronize());
  return NAX_OK;
}

extern "C" int nax_exchange(nax_sampler* h, float* z_out, const double* pe_in, const float* g_in) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  const EngineCfg& e = h->cfg;
  if (z_out)
    NAX_HIP(hipMemcpy2DAsync(z_out, sizeof(float) * e.D, h->zreq, sizeof(float) * e.Dp,
                             sizeof(float) * e.D, e.C, hipMemcpyDeviceToDevice, h->stream));
  if (pe_in)
    NAX_HIP(hipMemcpyAsync(h->peout, pe_in, sizeof(double) * e.C, hipMemcpyDeviceToDevice, h->stream));
  if (g_in)
    NAX_HIP(hipMemcpy2DAsync(h->gout, sizeof(float) * e.Dp, g_in, sizeof(float) * e.D,
                             sizeof(float) * e.D, e.C, hipMemcpyDeviceToDevice, h->stream));
  NAX_HIP(hipStreamSynchronize(h->stream));
  return NAX_OK;
}

extern "C" int nax_set_progress(nax_sampler* h, void (*cb)(void*, int64_t, int32_t), void* user,
                                int64_t every_ticks) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  h->progress_cb = cb;
  h->progress_user = user;
  h->progress_every = every_ticks > 0 ? (every_ticks + kBatch - 1) / kBatch * kBatch : 0;
  return NAX_OK;
}
