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
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "nax_device.h"
#include "nax_engine.h"
#include "nax_internal.h"

namespace nax {

// ---- the tick kernel: each chain runs to its next gradient request -------------
template <bool DENSE>
__global__ void __launch_bounds__(256)
advance_kernel(EngineCfg cfg, Collect col, int64_t target, ChainState* __restrict__ states,
               float* __restrict__ vec, double* __restrict__ welf, float* __restrict__ zreq,
               const float* __restrict__ gout, const double* __restrict__ peout,
               int* __restrict__ pending, int* __restrict__ active,
               int64_t* __restrict__ grad_evals) {
  __shared__ ChainState sS[4];
  const int c = wave_chain(4);
  if (c >= cfg.C) return;
  const ChainCtx cx = make_ctx<DENSE>(cfg, c, vec, welf);
  ChainState& S = sS[threadIdx.x / NAX_WAVE];
  copy_state(S, states[c], cx.lane);
  int evals = 0;
  const bool need_grad = advance_chain(cfg, col, target, S, cx, c, zreq + (size_t)c * cfg.Dp,
                                       gout + (size_t)c * cfg.Dp, peout[c], evals);
  copy_state(states[c], S, cx.lane);
  if (cx.lane == 0) {
    if (need_grad) active[atomicAdd(pending, 1)] = c;  // compacted request list
    if (evals && grad_evals) atomicAdd((unsigned long long*)grad_evals, (unsigned long long)evals);
  }
}

// init_kernel (hmc.py:193-362) + warmup_adapter.init_fn (hmc_util.py:552-586)
__global__ void __launch_bounds__(256)
init_kernel(EngineCfg cfg, ChainState* __restrict__ states, float* __restrict__ vec,
            double* __restrict__ welf, float* __restrict__ zreq, const float* __restrict__ z0,
            const uint32_t* __restrict__ keys, const float* __restrict__ imm0) {
  const int c = wave_chain(4);
  if (c >= cfg.C) return;
  const ChainCtx cx = make_ctx(cfg, c, vec, welf);
  const bool part = cfg.partitionable != 0;
  ChainState S;
  memset(&S, 0, sizeof(S));
  Key k = Key{keys[2 * c], keys[2 * c + 1]};
  Key k_hmc = split_key(k, 0, 3, part);
  Key k_wa = split_key(k, 1, 3, part);
  Key k_mom = split_key(k, 2, 3, part);
  setK(S.adapt_key, split_key(k_wa, 0, 2, part));
  setK(S.fs_key, split_key(k_wa, 1, 2, part));  // rng_key_ss of hmc_util.py:565
  setK(S.rng_key, k_hmc);
  float* imm = cx.v(V_IMM);
  float* msq = cx.v(V_MSQRT);
  if (cx.mi) {
    // _initialize_mass_matrix, dense (hmc_util.py:490-507): imm0 is [D][D]
    // (identity when absent); factors as at a window end.
    const size_t ld = cx.Dp;
    for (int i = 0; i < cx.D; ++i)
      FOR_D(j) {
        const double a = imm0 ? (double)imm0[(size_t)i * cx.D + j] : (i == j ? 1.0 : 0.0);
        cx.m2[i * ld + j] = a;
        cx.mi[(size_t)j * ld + i] = (float)a;
      }
    wave_sync_mem();
    dense_factor(cx);
    dense_zero(cx);
    FOR_D(d) { imm[d] = 1.0f; msq[d] = 1.0f; }
  } else {
    FOR_D(d) {
      float m = imm0 ? imm0[d] : 1.0f;
      imm[d] = m;
      msq[d] = imm0 ? (float)(1.0 / sqrt((double)m)) : 1.0f;
    }
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

// Dense mass matrices: buf [C][D][D] row-major <-> the transposed per-chain
// blocks blk [C][Dp][Dp] (blk[c][j][i] = buf[c][i][j]).
__global__ void dense_io(int C, int D, int Dp, float* blk, float* buf, int to_blk) {
  const int c = blockIdx.z;
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C || i >= D || j >= D) return;
  float* b = blk + (size_t)c * Dp * Dp + (size_t)j * Dp + i;
  float* u = buf + ((size_t)c * D + i) * D + j;
  if (to_blk) *b = *u;
  else *u = *b;
}

// Welford state <-> host-view buffers: mean [C][D] and the diagonal m2
// [C][D] live in wel [C][3][Dp] (mean, m2, scratch), fp64.
__global__ void welford_io(int C, int D, int Dp, double* wel, double* mean, double* m2,
                           int to_dev) {
  const int c = blockIdx.y;
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C || d >= D) return;
  double* w = wel + (size_t)c * 3 * Dp;
  if (to_dev) {
    if (mean) w[d] = mean[(size_t)c * D + d];
    if (m2) w[Dp + d] = m2[(size_t)c * D + d];
  } else {
    if (mean) mean[(size_t)c * D + d] = w[d];
    if (m2) m2[(size_t)c * D + d] = w[Dp + d];
  }
}

// Dense Welford co-moments: buf [C][D][D] <-> dm2 [C][Dp][Dp] (same
// orientation, m2[i][j] = sum dpost_i dpre_j).
__global__ void dense_m2_io(int C, int D, int Dp, double* dm2, double* buf, int to_dev) {
  const int c = blockIdx.z;
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C || i >= D || j >= D) return;
  double* b = dm2 + (size_t)c * Dp * Dp + (size_t)i * Dp + j;
  double* u = buf + ((size_t)c * D + i) * D + j;
  if (to_dev) *b = *u;
  else *u = *b;
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
  int* active = nullptr;       // device [C]: chains that requested a gradient this tick
  bool graphs = true;          // replay tick batches as hipGraphs (nax_set_graphs)
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
  e.find_heuristic = c->find_heuristic_step_size != 0;
  e.dense = c->dense_mass != 0;
  if (e.dense && c->dim > 1024) {
    set_error("dense_mass supports dim <= 1024");
    return NAX_EINVAL;
  }
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
  NAX_HIP(hipMalloc(&h->wel, C * 3 * Dp * sizeof(double)));
  if (e.dense) {
    NAX_HIP(hipMalloc(&e.dimm_t, C * Dp * Dp * sizeof(float)));
    NAX_HIP(hipMalloc(&e.dmsq_t, C * Dp * Dp * sizeof(float)));
    NAX_HIP(hipMalloc(&e.dm2, C * Dp * Dp * sizeof(double)));
    NAX_HIP(hipMemsetAsync(e.dimm_t, 0, C * Dp * Dp * sizeof(float), h->stream));
    NAX_HIP(hipMemsetAsync(e.dmsq_t, 0, C * Dp * Dp * sizeof(float), h->stream));
    NAX_HIP(hipMemsetAsync(e.dm2, 0, C * Dp * Dp * sizeof(double), h->stream));
  }
  NAX_HIP(hipMalloc(&h->st, C * sizeof(ChainState)));
  NAX_HIP(hipMalloc(&h->zreq, C * Dp * sizeof(float)));
  NAX_HIP(hipMalloc(&h->gout, C * Dp * sizeof(float)));
  NAX_HIP(hipMalloc(&h->peout, C * sizeof(double)));
  NAX_HIP(hipMalloc(&h->ctr, kBatch * sizeof(int)));
  NAX_HIP(hipMalloc(&h->active, C * sizeof(int)));
  NAX_HIP(hipMalloc(&h->gevals, 2 * sizeof(int64_t)));  // grad evals, persistent target
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
  hipFree(h->peout); hipFree(h->ctr); hipFree(h->gevals); hipFree(h->active);
  hipFree(h->cfg.dimm_t); hipFree(h->cfg.dmsq_t); hipFree(h->cfg.dm2);
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
  if (kernel < 0 || kernel > 4) {
    set_error("kernel must be 0 (auto), 1 (tile), 2 (matrix cores), 3 (one wavefront per chain) "
              "or 4 (single-copy bf16)");
    return NAX_EINVAL;
  }
  Potential* p = make_logistic(x, y, n, d, prior_scale, x_dtype, kernel, h->cfg.C, h->stream, &err);
  if (!p) { set_error(err); return NAX_EINVAL; }
  delete h->pot;
  h->pot = p;
  h->external = false;
  h->cfg.chain_dim = nullptr;
  return NAX_OK;
}

extern "C" int nax_model_sv_ex(nax_sampler* h, const float* returns, int32_t T,
                               int32_t noncentred, double sigma_rate, double nu_rate) {
  if (!h || !returns) { set_error("nax_model_sv: null argument"); return NAX_EINVAL; }
  if (T + 2 != h->cfg.D) { set_error("sv model: dim must be T + 2"); return NAX_EINVAL; }
  std::string err;
  Potential* p = make_sv(returns, T, noncentred, sigma_rate, nu_rate, h->stream, &err);
  if (!p) { set_error(err); return NAX_EINVAL; }
  delete h->pot;
  h->pot = p;
  h->external = false;
  h->cfg.chain_dim = nullptr;
  return NAX_OK;
}

extern "C" int nax_model_sv(nax_sampler* h, const float* returns, int32_t T, int32_t noncentred) {
  return nax_model_sv_ex(h, returns, T, noncentred, 50.0, 0.1);
}

extern "C" int nax_model_hier(nax_sampler* h, const float* x, const float* y,
                              const int64_t* row_off, const int64_t* elt_off, const int32_t* dims,
                              int32_t groups, int32_t chains_per_group) {
  if (!h || !x || !y || !row_off || !elt_off || !dims) {
    set_error("nax_model_hier: null argument");
    return NAX_EINVAL;
  }
  if (groups <= 0 || chains_per_group <= 0 || (int64_t)groups * chains_per_group != h->cfg.C) {
    set_error("hierarchical model: num_chains must equal groups * chains_per_group");
    return NAX_EINVAL;
  }
  std::string err;
  Potential* p = make_hier(x, y, row_off, elt_off, dims, groups, chains_per_group, h->cfg.C,
                           h->cfg.Dp, h->stream, &err);
  if (!p) { set_error(err); return NAX_EINVAL; }
  if (p->dim() != h->cfg.D) {
    delete p;
    set_error("hierarchical model: the sampler dim must be max(D_i) + 2");
    return NAX_EINVAL;
  }
  delete h->pot;
  h->pot = p;
  h->external = false;
  h->cfg.chain_dim = p->chain_dims();
  return NAX_OK;
}

extern "C" int nax_model_logistic_ecs(nax_sampler* h, const float* x, const float* y, int64_t n,
                                      int32_t d, double prior_scale, int32_t subsample_size,
                                      int32_t num_blocks, int32_t proxy_degree, const double* ref,
                                      int32_t* idx, uint32_t* keys) {
  if (!h || !x || !y) { set_error("nax_model_logistic_ecs: null argument"); return NAX_EINVAL; }
  if (d != h->cfg.D) { set_error("logistic model dim does not match the sampler dim"); return NAX_EINVAL; }
  std::string err;
  Potential* p = make_ecs_logistic(x, y, n, d, prior_scale, subsample_size, num_blocks,
                                   proxy_degree, ref, idx, keys, h->cfg.C, h->cfg.Dp,
                                   h->cfg.partitionable != 0, h->stream, &err);
  if (!p) { set_error(err); return NAX_EINVAL; }
  delete h->pot;
  h->pot = p;
  h->external = false;
  h->cfg.chain_dim = nullptr;
  return NAX_OK;
}

extern "C" int nax_ecs_gibbs(nax_sampler* h, float* accept_prob) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  if (!h->initialized) { set_error("nax_ecs_gibbs before nax_init"); return NAX_ESTATE; }
  if (!h->pot) { set_error("nax_ecs_gibbs: no subsampled model set"); return NAX_ESTATE; }
  const EngineCfg& e = h->cfg;
  hipError_t r = h->pot->ecs_gibbs(h->stream, h->st, h->vec + (size_t)V_Z * e.Dp,
                                   h->vec + (size_t)V_G * e.Dp, e.nvec * e.Dp, accept_prob);
  if (r == hipErrorNotSupported) {
    set_error("nax_ecs_gibbs: the model has no subsample plate (use nax_model_logistic_ecs)");
    return NAX_ESTATE;
  }
  NAX_HIP(r);
  return NAX_OK;
}

extern "C" int nax_model_external(nax_sampler* h) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  delete h->pot;
  h->pot = nullptr;
  h->external = true;
  return NAX_OK;
}

extern "C" int nax_set_launch_shape(nax_sampler* h, int32_t shape) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  if (!h->pot) { set_error("no built-in model set"); return NAX_ESTATE; }
  if (!h->pot->force_launch_shape(shape)) {
    set_error("nax_set_launch_shape: not a launch shape of this model (waves per chain 1, 2, 4, "
              "8, 16 for per-chain persistent models; a divisor of chains_per_group for the "
              "hierarchical model)");
    return NAX_EINVAL;
  }
  return NAX_OK;
}

extern "C" int nax_tuning(nax_sampler* h, char* buf, int32_t len) {
  if (!h || !buf || len < 1) { set_error("nax_tuning: null handle or buffer"); return NAX_EINVAL; }
  const std::string t = h->pot ? h->pot->tuning : std::string();
  const size_t n = std::min<size_t>(t.size(), (size_t)len - 1);
  memcpy(buf, t.data(), n);
  buf[n] = 0;
  return NAX_OK;
}

extern "C" int nax_set_positive_transform(nax_sampler* h, int32_t kind) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  if (kind != 0 && kind != 1) {
    set_error("positive transform: 0 (ExpTransform) or 1 (SoftplusTransform)");
    return NAX_EINVAL;
  }
  if (!h->pot) { set_error("no built-in model set"); return NAX_ESTATE; }
  if (h->initialized) {
    set_error("nax_set_positive_transform must come before nax_init");
    return NAX_ESTATE;
  }
  if (!h->pot->set_positive_transform(kind)) {
    set_error(std::string("model '") + h->pot->name() + "' has no positive latent site");
    return NAX_EINVAL;
  }
  return NAX_OK;
}

// Evaluates the potential for every chain's request.  With timing on, the
// main kernel of launch k of a batch is bracketed by events ev[2k], ev[2k+1];
// collect_timing() reads them after the batch has synchronised.
static int eval_events(nax_sampler* h, int k) {
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
  return NAX_OK;
}

// compact: evaluate only the chains the tick kernel of batch slot k listed.
static int eval_potential(nax_sampler* h, int k = 0, bool compact = false) {
  const EngineCfg& e = h->cfg;
  int rc = eval_events(h, k);
  if (rc) return rc;
  if (compact)
    NAX_HIP(h->pot->eval_active(h->stream, h->zreq, e.Dp, e.C, h->active, h->ctr + k, h->peout,
                                h->gout, e.Dp));
  else
    NAX_HIP(h->pot->eval(h->stream, h->zreq, e.Dp, e.C, h->peout, h->gout, e.Dp));
  h->pot->ev_begin = h->pot->ev_end = nullptr;
  return NAX_OK;
}

static int collect_timing(nax_sampler* h, int nlaunch, int first = 0) {
  if (!h->timing) return NAX_OK;
  for (int k = first; k < first + nlaunch; ++k) {
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
  if (e.dense)
    hipLaunchKernelGGL(advance_kernel<true>, grid, dim3(256), 0, h->stream, e, col, target,
                       h->st, h->vec, h->wel, h->zreq, h->gout, h->peout, h->ctr + slot, h->active,
                       h->gevals);
  else
    hipLaunchKernelGGL(advance_kernel<false>, grid, dim3(256), 0, h->stream, e, col, target,
                       h->st, h->vec, h->wel, h->zreq, h->gout, h->peout, h->ctr + slot, h->active,
                       h->gevals);
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
    // find_reasonable_step_size keeps chains asking for gradients until its
    // doubling / halving loop ends.
    for (;;) {
      NAX_HIP(hipMemcpyAsync(h->ctr_host, h->ctr, sizeof(int), hipMemcpyDeviceToHost, h->stream));
      NAX_HIP(hipStreamSynchronize(h->stream));
      if (h->ctr_host[0] == 0) break;
      rc = eval_potential(h, 0, true);
      if (rc) return rc;
      hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, h->stream, h->ctr, kBatch);
      rc = launch_advance(h, col, 0, 0);
      if (rc) return rc;
    }
  }
  return NAX_OK;
}

extern "C" int nax_run(nax_sampler* h, int64_t target, const nax_collect* c, int64_t* gevals) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  if (!h->initialized) { set_error("nax_run before nax_init"); return NAX_ESTATE; }
  if (h->external || !h->pot) { set_error("nax_run needs a built-in model; use nax_advance"); return NAX_ESTATE; }
  Collect col = to_collect(c);
  NAX_HIP(hipMemsetAsync(h->gevals, 0, sizeof(int64_t), h->stream));
  if (h->pot->persistent()) {
    // Whole loop on the device, in chunks of iterations with progress
    // reported between them.  Every chunk ends when its slowest chain does,
    // so chunks are long: the first is kPersistIters iterations, later ones
    // are sized from its duration to about kChunkMs of work (at most
    // kMaxChunk iterations).  The iteration target lives in device memory
    // (gevals[1]): a chunk is a target step kernel plus the persistent
    // launch, enqueued directly or replayed as a hipGraph captured for the
    // current chunk length (the model's autotuning runs before, uncaptured;
    // a model that tunes at iteration tune_at has a chunk end there and is
    // prepared again).  Kernel timing brackets the main kernel with events,
    // so it runs uncaptured.
    constexpr int64_t kPersistIters = 50, kMaxChunk = 1000;
    constexpr double kChunkMs = 1000.0;
    std::vector<ChainState> probe(1);
    NAX_HIP(hipMemcpyAsync(probe.data(), h->st, sizeof(ChainState), hipMemcpyDeviceToHost, h->stream));
    NAX_HIP(hipStreamSynchronize(h->stream));
    int64_t* d_target = h->gevals + 1;
    int64_t t = probe[0].i;
    h->pot->chain_iter = t;
    NAX_HIP(h->pot->prepare_persistent(h->stream, h->cfg, h->st, h->vec, h->wel, h->zreq,
                                       h->gout, h->peout));
    hipLaunchKernelGGL(set_target_kernel, dim3(1), dim3(1), 0, h->stream, d_target, t);
    NAX_HIP(hipGetLastError());
    int64_t step = kPersistIters;
    auto chunk = [&](int64_t len) -> hipError_t {
      hipLaunchKernelGGL(step_target_kernel, dim3(1), dim3(1), 0, h->stream, d_target, target,
                         len);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      return h->pot->run_persistent(h->stream, h->cfg, col, d_target, h->st, h->vec, h->wel,
                                    h->zreq, h->gout, h->peout, h->gevals);
    };
    const bool use_graph = h->graphs && !h->timing;
    hipGraphExec_t exec = nullptr;
    struct ExecGuard {
      hipGraphExec_t* e;
      ~ExecGuard() { if (*e) hipGraphExecDestroy(*e); }
    } guard{&exec};
    int64_t launches = 0;
    while (t < target) {
      // a chunk ends at the model's tuning iteration (run directly: its
      // length differs from the captured one)
      const int64_t ta = h->pot->tune_at;
      const int64_t len = t < ta ? std::min<int64_t>(step, ta - t) : step;
      const int64_t t0 = t;
      t = std::min<int64_t>(target, t + len);
      const int k = (int)(launches % kBatch);
      const auto wall0 = std::chrono::steady_clock::now();
      if (use_graph && len == step) {
        if (!exec) {
          NAX_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
          const hipError_t le = chunk(step);
          hipGraph_t graph = nullptr;
          const hipError_t ce = hipStreamEndCapture(h->stream, &graph);
          if (le != hipSuccess || ce != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            NAX_HIP(le != hipSuccess ? le : ce);
          }
          const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
          hipGraphDestroy(graph);
          NAX_HIP(ie);
        }
        NAX_HIP(hipGraphLaunch(exec, h->stream));
      } else if (use_graph) {
        NAX_HIP(chunk(len));
      } else {
        int rc = eval_events(h, k);
        if (rc) return rc;
        NAX_HIP(chunk(len));
        h->pot->ev_begin = h->pot->ev_end = nullptr;
      }
      NAX_HIP(hipStreamSynchronize(h->stream));
      if (!use_graph) {
        const int rc = collect_timing(h, 1, k);
        if (rc) return rc;
      }
      if (launches == 0 && t < target) {
        const double ms = std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - wall0).count();
        const int64_t next = std::max<int64_t>(
            kPersistIters, std::min<int64_t>(kMaxChunk, (int64_t)(len * kChunkMs /
                                                                   std::max(ms, 1e-3))));
        if (next != step) {
          step = next;
          if (exec) {  // captured with the old chunk length
            hipGraphExecDestroy(exec);
            exec = nullptr;
          }
        }
      }
      if (t0 < ta && t >= ta && t < target) {
        h->pot->chain_iter = t;
        const std::string before = h->pot->tuning;
        NAX_HIP(h->pot->prepare_persistent(h->stream, h->cfg, h->st, h->vec, h->wel, h->zreq,
                                           h->gout, h->peout));
        if (exec && h->pot->tuning != before) {  // captured with the old launch shape
          hipGraphExecDestroy(exec);
          exec = nullptr;
        }
      }
      ++launches;
      if (h->progress_cb) h->progress_cb(h->progress_user, launches, h->cfg.C);
    }
    if (gevals) {
      NAX_HIP(hipMemcpyAsync(h->gevals_host, h->gevals, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
      NAX_HIP(hipStreamSynchronize(h->stream));
      *gevals = h->gevals_host[0];
    }
    return NAX_OK;
  }
  // One batch = kBatch ticks (tick kernel + potential), enqueued directly or
  // replayed as a hipGraph captured after two direct batches (so workspaces
  // are allocated and short calls never pay for a capture).  The graph holds
  // kernels only (counters are cleared by a kernel, not a memset).  Kernel
  // timing records events around the main kernel, so it runs uncaptured.
  auto batch = [&]() -> int {
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, h->stream, h->ctr, kBatch);
    for (int k = 0; k < kBatch; ++k) {
      int rc = launch_advance(h, col, target, k);
      if (rc) return rc;
      rc = eval_potential(h, k, true);
      if (rc) return rc;
    }
    return NAX_OK;
  };
  const bool use_graph = h->graphs && !h->timing;
  hipGraphExec_t exec = nullptr;
  struct ExecGuard {
    hipGraphExec_t* e;
    ~ExecGuard() { if (*e) hipGraphExecDestroy(*e); }
  } guard{&exec};
  int64_t ticks = 0;
  for (;;) {
    if (use_graph && !exec && ticks >= 2 * kBatch) {
      NAX_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      const int rc = batch();
      hipGraph_t graph = nullptr;
      const hipError_t ce = hipStreamEndCapture(h->stream, &graph);
      if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
      }
      NAX_HIP(ce);
      const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      NAX_HIP(ie);
    }
    if (exec) {
      NAX_HIP(hipGraphLaunch(exec, h->stream));
    } else {
      const int rc = batch();
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
  if (e.dense) {
    // [C, dim, dim] matrices instead of the diagonal slots
    dim3 g2((e.D + 127) / 128, e.D, e.C);
    if (v->d_inverse_mass_matrix)
      hipLaunchKernelGGL(dense_io, g2, dim3(128), 0, h->stream, e.C, e.D, e.Dp, e.dimm_t,
                         v->d_inverse_mass_matrix, to_dev ? 1 : 0);
    if (v->d_mass_matrix_sqrt)
      hipLaunchKernelGGL(dense_io, g2, dim3(128), 0, h->stream, e.C, e.D, e.Dp, e.dmsq_t,
                         v->d_mass_matrix_sqrt, to_dev ? 1 : 0);
    NAX_HIP(hipGetLastError());
  }
  for (auto& x : vs) {
    if (!x.p) continue;
    if (e.dense && (x.slot == V_IMM || x.slot == V_MSQRT)) continue;
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

static int adapt_io(nax_sampler* h, const nax_adapt_view* v, bool to_dev) {
  const EngineCfg& e = h->cfg;
  if (v->d_mm_mean || (v->d_mm_m2 && !e.dense)) {
    hipLaunchKernelGGL(welford_io, dim3((e.D + 127) / 128, e.C), dim3(128), 0, h->stream, e.C,
                       e.D, e.Dp, h->wel, v->d_mm_mean, e.dense ? nullptr : v->d_mm_m2,
                       to_dev ? 1 : 0);
    NAX_HIP(hipGetLastError());
  }
  if (v->d_mm_m2 && e.dense) {
    hipLaunchKernelGGL(dense_m2_io, dim3((e.D + 127) / 128, e.D, e.C), dim3(128), 0, h->stream,
                       e.C, e.D, e.Dp, e.dm2, v->d_mm_m2, to_dev ? 1 : 0);
    NAX_HIP(hipGetLastError());
  }
  std::vector<ChainState> hs(e.C);
  NAX_HIP(hipMemcpyAsync(hs.data(), h->st, sizeof(ChainState) * e.C, hipMemcpyDeviceToHost,
                         h->stream));
  NAX_HIP(hipStreamSynchronize(h->stream));
  for (int c = 0; c < e.C; ++c) {
    ChainState& s = hs[c];
    if (to_dev) {
      if (v->ss_x_t) s.da_x = v->ss_x_t[c];
      if (v->ss_x_avg) s.da_xavg = v->ss_x_avg[c];
      if (v->ss_g_avg) s.da_gavg = v->ss_g_avg[c];
      if (v->ss_t) s.da_t = v->ss_t[c];
      if (v->ss_prox_center) s.da_prox = v->ss_prox_center[c];
      if (v->mm_n) s.welford_n = v->mm_n[c];
      if (v->window_idx) s.window_idx = v->window_idx[c];
      if (v->rng_key) { s.adapt_key[0] = v->rng_key[2 * c]; s.adapt_key[1] = v->rng_key[2 * c + 1]; }
    } else {
      if (v->ss_x_t) v->ss_x_t[c] = s.da_x;
      if (v->ss_x_avg) v->ss_x_avg[c] = s.da_xavg;
      if (v->ss_g_avg) v->ss_g_avg[c] = s.da_gavg;
      if (v->ss_t) v->ss_t[c] = s.da_t;
      if (v->ss_prox_center) v->ss_prox_center[c] = s.da_prox;
      if (v->mm_n) v->mm_n[c] = s.welford_n;
      if (v->window_idx) v->window_idx[c] = s.window_idx;
      if (v->rng_key) { v->rng_key[2 * c] = s.adapt_key[0]; v->rng_key[2 * c + 1] = s.adapt_key[1]; }
    }
  }
  if (to_dev) {
    for (int c = 0; c < e.C; ++c)
      if (hs[c].window_idx < 0 || hs[c].window_idx > e.num_windows || hs[c].welford_n < 0 ||
          hs[c].da_t < 0) {
        set_error("nax_set_adapt_state: window_idx, Welford count or dual-averaging t out of "
                  "range");
        return NAX_EINVAL;
      }
    NAX_HIP(hipMemcpyAsync(h->st, hs.data(), sizeof(ChainState) * e.C, hipMemcpyHostToDevice,
                           h->stream));
  }
  NAX_HIP(hipStreamSynchronize(h->stream));
  return NAX_OK;
}

extern "C" int nax_get_adapt_state(nax_sampler* h, const nax_adapt_view* v) {
  if (!h || !v) { set_error("null argument"); return NAX_EINVAL; }
  if (!h->initialized) { set_error("no state before nax_init"); return NAX_ESTATE; }
  return adapt_io(h, v, false);
}

extern "C" int nax_set_adapt_state(nax_sampler* h, const nax_adapt_view* v) {
  if (!h || !v) { set_error("null argument"); return NAX_EINVAL; }
  if (!h->initialized) { set_error("nax_set_adapt_state before nax_init"); return NAX_ESTATE; }
  return adapt_io(h, v, true);
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

extern "C" int nax_set_graphs(nax_sampler* h, int32_t enable) {
  if (!h) { set_error("null handle"); return NAX_EINVAL; }
  h->graphs = enable != 0;
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
