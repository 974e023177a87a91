// Many independent non-centred hierarchical logistic regressions of uneven
// size, K chains per model, in one launch.
//
// Model i (rows N_i, columns D_i <= 64):
//   eta ~ Normal(0, 1)^D_i, mu ~ Normal(0, 1), tau ~ HalfCauchy(1),
//   beta = mu + tau * eta, y_n ~ Bernoulli(logits = x_n . beta)
// (Normal continuous.py:2975-2989, HalfCauchy continuous.py:1149-1150 ->
// Cauchy continuous.py:390-406, BernoulliLogits discrete.py:244-263; tau is
// unconstrained by ExpTransform, transforms.py:606-652, log|J| = log tau).
// Latent of a chain, sorted site names: [eta (D_i), mu, log tau].
//
// Data are ragged row-major blocks: X rows of model i at elt_off[i], labels at
// row_off[i].  A block of K wavefronts owns model i and its K chains: one X
// tile of 64 rows is staged in LDS once and used by all K chains (phase 1:
// lane = row, wave = chain -> logits and residuals; phase 2: lane = column,
// wave = chain -> X^T R), so each byte of X_i is read once per gradient tick
// of the model, not once per chain.  In the persistent form each wave runs
// its chain's NUTS state machine; when any chain of the block needs a
// gradient the block evaluates all K together (chains that are already done
// ride along and ignore the result).  Blocks take models heaviest first
// (N_i D_i descending) to even out the tail of the launch.  A model's K
// chains may also be split over K / KB workgroups of KB chains each (each
// then reads X_i for its KB chains): more workgroups for the heavy models
// against more passes over their X.  And each chain may own R waves: the
// block then stages R tiles (64 R rows) at a time, wave r of a chain takes
// rows 64 r .. 64 r + 63 of them in both phases, and the R partial sums of
// U and X^T R are added in r order at the end -- a heavy model's leaf then
// takes ~1/R of the time, which is what the launch waits for.  The NUTS
// state machine of a chain runs on its wave r = 0.  The autotuner picks
// (KB, R).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "nax_device.h"
#include "nax_engine.h"
#include "nax_internal.h"

namespace nax {

constexpr int kHierRows = 64;
constexpr int kHierMaxD = 64;
constexpr double kLog2PiHalf = 0.91893853320467274;
constexpr double kLogPi_ = 1.1447298858494002;
constexpr double kLog2 = 0.69314718055994531;

struct HierData {
  const float* x;
  const float* y;
  const int64_t* row_off;
  const int64_t* elt_off;
  const int* dims;
  const int* order;  // block -> model, heaviest first
  int K;             // chains per model
  int Dp;            // stride of a chain's latent vector
  int softplus;      // tau by SoftplusTransform instead of ExpTransform
};

// Row stride of a staged X tile: 68 floats keeps both access patterns free
// of LDS bank conflicts (lane = row reading 16-byte column groups, lane =
// column reading one row) and 16-byte aligned.
constexpr int kHierLd = kHierMaxD + 4;

template <int K, int R>
struct __attribute__((aligned(16))) HierShared {
  __attribute__((aligned(16))) float xs[2][kHierRows * R][kHierLd];  // double-buffered X tiles
  __attribute__((aligned(16))) float beta[K][kHierMaxD];
  __attribute__((aligned(16))) float res[K][kHierRows * R];  // residuals sigmoid(l) - y, chain-major
  float ys[2][kHierRows * R];
  // partial X^T R and U of the R waves of each chain (R > 1)
  __attribute__((aligned(16))) float racc[R > 1 ? K : 1][R > 1 ? R : 1][kHierMaxD];
  double rul[R > 1 ? K : 1][R > 1 ? R : 1];
};

// Tiles held in registers ahead of the one being processed: enough bytes in
// flight per workgroup to cover the memory latency when a model's workgroup
// runs alone on its CU (the tail of a launch), within ~64 VGPRs.
template <int K, int R>
struct HierPipe {
  static constexpr int NT = 64 * K * R;                       // threads
  static constexpr int EPT = kHierRows * R * kHierMaxD / NT;  // tile elements per thread
  static constexpr int PF = (EPT >= 32) ? 1 : 2;
};

// Block-cooperative U and dU for the K chains of model i at z[c] (c = chain
// index of the block's chain slot w: c = i K_model + sub K + w), each chain
// on R waves (wave index w R + r).  Wave r = 0 of slot w ends with its own
// chain's results: returns U on every lane of that wave (0 on the others),
// writes dU to g[c].  Tiles of 64 R rows are contiguous (rows * D floats) in
// the ragged layout.  Tile t is loaded into registers PF tiles ahead, staged
// into LDS buffer t & 1, then: phase 1 (lane = row 64 r + lane, wave =
// chain part) logits and residuals, phase 2 (lane = column, same wave) X^T R
// over the same 64 rows.  One barrier per tile: the residuals a wave writes
// in phase 1 are read back only by that wave (LDS accesses of one wavefront
// complete in order), and buffer t & 1 is next written (tile t + 2) after
// every wave has passed the barrier of tile t + 1.  With R > 1 the R
// partial sums of X^T R and of U are added in r order after the last tile.
template <int K, int R>
__device__ double hier_group_eval(const HierData& hd, int i, int sub, const float* zreq,
                                  int ldz, float* gout, int ldg, HierShared<K, R>& sh) {
  using P = HierPipe<K, R>;
  constexpr int NT = P::NT, EPT = P::EPT, PF = P::PF;
  constexpr int TR = kHierRows * R;  // rows of a staged tile
  const int wa = threadIdx.x >> 6;
  const int w = wa / R;               // chain slot
  const int r = wa % R;               // 64-row part of the tile
  const int lane = threadIdx.x & 63;
  const int D = hd.dims[i];
  const int D4 = (D + 3) & ~3;
  const int64_t N = hd.row_off[i + 1] - hd.row_off[i];
  const int ntiles = (int)((N + TR - 1) / TR);
  const float* X = hd.x + hd.elt_off[i];
  const float* Y = hd.y + hd.row_off[i];
  const int c = i * hd.K + sub * K + w;
  const float* z = zreq + (size_t)c * ldz;
  float* g = gout + (size_t)c * ldg;
  const float mu = z[D];
  const PosTf tt = positive_tf((double)z[D + 1], hd.softplus != 0);
  const float tau = hd.softplus ? (float)tt.x : expf(z[D + 1]);
  const float eta = lane < D ? z[lane] : 0.0f;
  if (r == 0) sh.beta[w][lane] = lane < D ? mu + tau * eta : 0.0f;
  // columns [D, D4) of both buffers stay zero for the whole evaluation
  for (int e = threadIdx.x; e < 2 * TR * 4; e += NT) {
    const int b = e / (TR * 4), rr = (e / 4) % TR, cc = D + (e & 3);
    if (cc < D4) sh.xs[b][rr][cc] = 0.0f;
  }
  // element e = threadIdx.x + NT j of a tile sits at row e / D, column e % D
  const int step_r = NT / D, step_c = NT % D;
  const int r_first = threadIdx.x / D, c_first = threadIdx.x % D;
  float px[PF][EPT];
  float py[PF];
  auto load = [&](int t, float* dst, float& ydst) {
    const int64_t r0 = (int64_t)t * TR;
    const int rows = (int)min<int64_t>(TR, N - r0);
    const int ne = rows * D;
    const float* src = X + r0 * D;
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int e = threadIdx.x + NT * j;
      dst[j] = e < ne ? src[e] : 0.0f;
    }
    ydst = threadIdx.x < rows ? Y[r0 + threadIdx.x] : 0.0f;
  };
#pragma unroll
  for (int j = 0; j < PF; ++j)
    if (j < ntiles) load(j, px[j], py[j]);
  const int rr0 = kHierRows * r;  // this wave's first row in a tile
  float acc = 0.0f;   // phase 2: dU/dbeta_lane for chain w over this wave's rows
  double ul = 0.0;    // phase 1: likelihood part of U for chain w over this wave's rows
  for (int t0 = 0; t0 < ntiles; t0 += PF) {
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      const int t = t0 + j;
      if (t < ntiles) {
        const int rows = (int)min<int64_t>(TR, N - (int64_t)t * TR);
        float(*xs)[kHierLd] = sh.xs[t & 1];
        {
          // the whole tile is written: rows past `rows` get zeros
          const int ne = TR * D;
          int rr = r_first, cc = c_first;
#pragma unroll
          for (int q = 0; q < EPT; ++q) {
            if (threadIdx.x + NT * q < ne) xs[rr][cc] = px[j][q];
            rr += step_r;
            cc += step_c;
            if (cc >= D) { cc -= D; rr += 1; }
          }
          if (threadIdx.x < TR) sh.ys[t & 1][threadIdx.x] = py[j];
        }
        if (t + PF < ntiles) load(t + PF, px[j], py[j]);  // in flight during the compute below
        __syncthreads();
        const int rows_r = min(max(rows - rr0, 0), kHierRows);
        float rv = 0.0f;
        if (lane < rows_r) {
          float l0 = 0.0f, l1 = 0.0f, l2 = 0.0f, l3 = 0.0f;
          const float4* xr = reinterpret_cast<const float4*>(&xs[rr0 + lane][0]);
          const float4* br = reinterpret_cast<const float4*>(&sh.beta[w][0]);
#pragma unroll
          for (int q = 0; q < kHierMaxD / 4; ++q) {
            if (4 * q < D4) {
              const float4 xv = xr[q], bv = br[q];
              l0 = fmaf(xv.x, bv.x, l0);
              l1 = fmaf(xv.y, bv.y, l1);
              l2 = fmaf(xv.z, bv.z, l2);
              l3 = fmaf(xv.w, bv.w, l3);
            }
          }
          float u;
          logistic_terms_f((l0 + l1) + (l2 + l3), sh.ys[t & 1][rr0 + lane], rv, u);
          ul += (double)u;
        }
        sh.res[w][rr0 + lane] = rv;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < D) {
          const int rows4 = (rows_r + 3) & ~3;  // residuals of rows past `rows` are 0
          const float4* rq = reinterpret_cast<const float4*>(&sh.res[w][rr0]);
          float a0 = 0.0f, a1 = 0.0f;
          for (int q = 0; q < rows4 / 4; ++q) {
            const float4 rv4 = rq[q];
            a0 = fmaf(xs[rr0 + 4 * q][lane], rv4.x, a0);
            a1 = fmaf(xs[rr0 + 4 * q + 1][lane], rv4.y, a1);
            a0 = fmaf(xs[rr0 + 4 * q + 2][lane], rv4.z, a0);
            a1 = fmaf(xs[rr0 + 4 * q + 3][lane], rv4.w, a1);
          }
          acc += a0 + a1;
        }
      }
    }
  }
  double U_lik;
  if constexpr (R > 1) {
    sh.racc[w][r][lane] = acc;
    const double uw = wave_sum(ul);
    if (lane == 0) sh.rul[w][r] = uw;
    __syncthreads();
    if (r != 0) return 0.0;
    float a = 0.0f;
    double us = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      a += sh.racc[w][q][lane];
      us += sh.rul[w][q];
    }
    acc = a;
    U_lik = us;
  } else {
    U_lik = wave_sum(ul);
  }
  // chain w's own gradient: dU/deta = eta + tau acc, dU/dmu = mu + sum acc,
  // dU/dlog tau = tau eta.acc + 2 tau^2 / (1 + tau^2) - 1
  const double sum_acc = wave_sum((double)acc);
  const double eta_acc = wave_sum((double)eta * acc);
  const double eta2 = wave_sum((double)eta * eta);
  if (lane < D) g[lane] = eta + tau * acc;
  const double t2 = (double)tau * tau;
  if (lane == 0) {
    g[D] = (float)((double)mu + sum_acc);
    // softplus: dtau/du (eta.acc + 2 tau / (1 + tau^2)) - d log|J| / du
    g[D + 1] = hd.softplus
        ? (float)(tt.dx * (eta_acc + 2.0 * (double)tau / (1.0 + t2)) - tt.dladj)
        : (float)((double)tau * eta_acc + 2.0 * t2 / (1.0 + t2) - 1.0);
  }
  const double U = U_lik + 0.5 * eta2 + D * kLog2PiHalf          // eta ~ N(0,1)
                   + 0.5 * (double)mu * mu + kLog2PiHalf          // mu ~ N(0,1)
                   - (kLog2 - kLogPi_ - log1p(t2))                // HalfCauchy(1)
                   - (hd.softplus ? tt.ladj : (double)z[D + 1]);  // log|J|
  return U;
}

template <int K, int R>
__global__ void __launch_bounds__(64 * K * R)
hier_eval_kernel(HierData hd, const float* __restrict__ zreq, int ldz, double* __restrict__ pe,
                 float* __restrict__ gout, int ldg) {
  __shared__ HierShared<K, R> sh;
  const int per = hd.K / K;  // workgroups per model
  const int i = hd.order[blockIdx.x / per];
  const int sub = blockIdx.x % per;
  const double u = hier_group_eval<K, R>(hd, i, sub, zreq, ldz, gout, ldg, sh);
  const int wa = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0 && wa % R == 0) pe[i * hd.K + sub * K + wa / R] = u;
}

// One block per (model, chain subset): K chain slots of R waves each.  Wave
// r = 0 of a slot runs its chain's NUTS state machine; every wave of the
// block joins each gradient evaluation.
template <int K, int R, bool DENSE>
__global__ void __launch_bounds__(64 * K * R, (K * R <= 8) ? 2 : 1)
hier_persistent_kernel(HierData hd, EngineCfg cfg, Collect col,
                       const int64_t* __restrict__ d_target, ChainState* __restrict__ states,
                       float* __restrict__ vec, double* __restrict__ welf,
                       float* __restrict__ zreq, float* __restrict__ gout,
                       double* __restrict__ peout, int64_t* __restrict__ grad_evals) {
  __shared__ HierShared<K, R> sh;
  __shared__ int need_any;
  __shared__ ChainState sS[K];
  const int64_t target = *d_target;
  const int per = hd.K / K;  // workgroups per model
  const int i = hd.order[blockIdx.x / per];
  const int sub = blockIdx.x % per;
  const int wa = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int w = wa / R;
  const bool head = wa % R == 0;  // the wave that runs the state machine
  const int c = i * hd.K + sub * K + w;
  const ChainCtx cx = make_ctx<DENSE>(cfg, c, vec, welf);
  float* zr = zreq + (size_t)c * cfg.Dp;
  float* gi = gout + (size_t)c * cfg.Dp;
  ChainState& S = sS[w];
  if (head) copy_state(S, states[c], cx.lane);
  double pe = peout[c];
  int evals = 0;
  bool active = head;
  for (;;) {
    if (threadIdx.x == 0) need_any = 0;
    __syncthreads();
    bool need = false;
    if (active) {
      need = advance_chain(cfg, col, target, S, cx, c, zr, gi, pe, evals);
      active = need;
    }
    if (need && cx.lane == 0) atomicOr(&need_any, 1);
    __threadfence_block();
    __syncthreads();
    if (!need_any) break;
    const double u = hier_group_eval<K, R>(hd, i, sub, zreq, cfg.Dp, gout, cfg.Dp, sh);
    if (need) pe = u;
    __threadfence_block();
  }
  if (head) {
    copy_state(states[c], S, cx.lane);
    if (cx.lane == 0) {
      peout[c] = pe;
      if (evals && grad_evals) atomicAdd((unsigned long long*)grad_evals, (unsigned long long)evals);
    }
  }
}

__global__ void hier_chain_dims(const int* dims, int K, int C, int* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C) out[c] = dims[c / K] + 2;
}

struct HierPotential : Potential {
  HierData hd;
  int groups = 0, C = 0, maxd = 0;
  int* order = nullptr;
  int* dims = nullptr;
  int* cdims = nullptr;
  int64_t* row_off = nullptr;
  int64_t* elt_off = nullptr;
  double flops = 0.0, bytes = 0.0;

  ~HierPotential() override {
    hipFree(order); hipFree(dims); hipFree(cdims); hipFree(row_off); hipFree(elt_off);
  }
  int dim() const override { return maxd + 2; }
  const char* name() const override { return "hier"; }
  bool set_positive_transform(int kind) override {
    hd.softplus = kind == 1;
    return true;
  }
  bool persistent() const override { return true; }
  const int* chain_dims() const override { return cdims; }
  double alg_bytes(int) const override { return bytes; }
  double alg_flops(int) const override { return flops; }

  // launch shape: kb chains per workgroup (a divisor of hd.K; 0: hd.K) and
  // rw waves per chain, code = kb + 32 (rw - 1).  The default (hd.K, 1) runs
  // until iteration kHierTuneAt, where the tuned shape (kb_t, rw_t) takes
  // over -- timed on trees of working size, not the first iterations' few
  // leaves -- unless a shape was fixed.
  static constexpr int64_t kHierTuneAt = 50;
  int kb = 0, rw = 1;
  int kb_t = 0, rw_t = 1;
  std::string tuned_log;
  int tuned_chains = 0;
  bool forced = false;
  HierPotential() { tune_at = kHierTuneAt; }
  static int shape_code(int k, int r) { return k + 32 * (r - 1); }
  static std::string shape_label(int code) {
    return std::to_string(code % 32) + "x" + std::to_string(code / 32 + 1);
  }
  // the (kb, rw) pairs with a kernel: rw in {1, 2, 4}, kb rw <= 16
  static bool has_kernel(int k, int r) {
    return (r == 1 && (k == 1 || k == 2 || k == 4 || k == 8 || k == 16)) ||
           ((r == 2 || r == 4) && (k == 1 || k == 2 || k == 4) && k * r <= 16);
  }
  void describe(const std::string& how) {
    tuning = "hier: " + std::to_string(waves()) + " chains per workgroup of " +
             std::to_string(hd.K) + " per model, " + std::to_string(rw) +
             " waves per chain " + how;
  }
  bool force_launch_shape(int shape) override {
    const int k = shape % 32, r = shape / 32 + 1;
    if (k < 1 || k > hd.K || hd.K % k || !has_kernel(k, r)) return false;
    kb = k;
    rw = r;
    forced = true;
    tune_at = 0;
    describe("(fixed)");
    return true;
  }

  int waves() const { return kb > 0 ? kb : hd.K; }

  hipError_t eval(hipStream_t s, const float* z, int ldz, int Cn, double* pe, float* g,
                  int ldg) override {
    if (Cn != C || ldz < maxd + 2 || ldg < maxd + 2) return hipErrorInvalidValue;
    // one evaluation of every chain: one wave per chain unless a shape was
    // fixed (C4 at fixed positions: 0.224 ms with one wave per chain, 0.299
    // with two, 0.255 with four -- the waves per chain pay off only on the
    // heavy models' chains that the persistent launch waits for)
    const int w = waves(), r = forced ? rw : 1;
    const dim3 grid(groups * (hd.K / w));
    bool launched = false;
    mark_begin(s);
#define NAX_KR(KV, RV)                                                                        \
  if (!launched && w == KV && r == RV) {                                                      \
    hipLaunchKernelGGL((hier_eval_kernel<KV, RV>), grid, dim3(64 * KV * RV), 0, s, hd, z,     \
                       ldz, pe, g, ldg);                                                      \
    launched = true;                                                                          \
  }
    NAX_KR(1, 1) NAX_KR(2, 1) NAX_KR(4, 1) NAX_KR(8, 1) NAX_KR(16, 1)
    NAX_KR(1, 2) NAX_KR(2, 2) NAX_KR(4, 2)
    NAX_KR(1, 4) NAX_KR(2, 4) NAX_KR(4, 4)
#undef NAX_KR
    if (!launched) return hipErrorInvalidValue;
    mark_end(s);
    return hipGetLastError();
  }

  hipError_t launch(int code, hipStream_t s, const EngineCfg& cfg, const Collect& col,
                    const int64_t* d_target, ChainState* st, float* vec, double* wel, float* zreq,
                    float* gout, double* peout, int64_t* gevals) const {
    const int w = code % 32, r = code / 32 + 1;
    if (w < 1 || hd.K % w) return hipErrorInvalidValue;
    const dim3 grid(groups * (hd.K / w));
#define NAX_KR(KV, RV)                                                                        \
  if (w == KV && r == RV) {                                                                   \
    if (cfg.dense)                                                                            \
      hipLaunchKernelGGL((hier_persistent_kernel<KV, RV, true>), grid, dim3(64 * KV * RV), 0,  \
                         s, hd, cfg, col, d_target, st, vec, wel, zreq, gout, peout, gevals); \
    else                                                                                      \
      hipLaunchKernelGGL((hier_persistent_kernel<KV, RV, false>), grid, dim3(64 * KV * RV), 0, \
                         s, hd, cfg, col, d_target, st, vec, wel, zreq, gout, peout, gevals); \
    return hipGetLastError();                                                                 \
  }
    NAX_KR(1, 1) NAX_KR(2, 1) NAX_KR(4, 1) NAX_KR(8, 1) NAX_KR(16, 1)
    NAX_KR(1, 2) NAX_KR(2, 2) NAX_KR(4, 2)
    NAX_KR(1, 4) NAX_KR(2, 4) NAX_KR(4, 4)
#undef NAX_KR
    return hipErrorInvalidValue;
  }

  hipError_t prepare_persistent(hipStream_t s, const EngineCfg& cfg, ChainState* st, float* vec,
                                double* wel, float* zreq, float* gout, double* peout) override {
    // candidates: one workgroup per model (kb = K) with 1, 2 or 4 waves per
    // chain, or each model's chains split over 2 or 4 workgroups (K/2, K/4)
    // with one wave each; timed over 5 iterations from a scratch copy of
    // the state.  The choice is kept per process for this shape of workload
    // (the waves per chain set the order of the sums, so the draws).
    if (forced) return hipSuccess;
    if (chain_iter < tune_at) {
      kb = 0;
      rw = 1;
      describe("(default until iteration " + std::to_string(tune_at) + ", then tuned)");
      return hipSuccess;
    }
    if (kb_t == 0 || tuned_chains != cfg.C) {
      const std::string key = "hier G=" + std::to_string(groups) + " K=" + std::to_string(hd.K) +
                              " C=" + std::to_string(cfg.C) + " dense=" + std::to_string(cfg.dense);
      int code = 0;
      if (tuned_shape(key, &code, false)) {
        tuned_log = "(tuned earlier in this process)";
      } else {
        std::vector<int> cands;
        for (int r : {1, 2, 4})
          if (has_kernel(hd.K, r)) cands.push_back(shape_code(hd.K, r));
        for (int w = hd.K / 2; w >= 1 && w * 4 >= hd.K; w /= 2)
          if (has_kernel(w, 1)) cands.push_back(shape_code(w, 1));
        std::string log;
        auto go = [&](int c, hipStream_t ss, const Collect& col, const int64_t* tgt,
                      ChainState* st2, float* vec2, double* wel2, float* zr2, float* g2,
                      double* pe2, int64_t* ge2) {
          return launch(c, ss, cfg, col, tgt, st2, vec2, wel2, zr2, g2, pe2, ge2);
        };
        code = tune_persistent(s, cfg, st, vec, wel, zreq, gout, peout, 5, cands, go, &log,
                               &HierPotential::shape_label);
        tuned_shape(key, &code, true);
        tuned_log = "(from iteration " + std::to_string(tune_at) +
                    "; ms for 5 iterations, chains per workgroup x waves per chain:" + log + ")";
      }
      kb_t = code % 32;
      rw_t = code / 32 + 1;
      tuned_chains = cfg.C;
    }
    kb = kb_t;
    rw = rw_t;
    describe(tuned_log);
    return hipSuccess;
  }

  hipError_t run_persistent(hipStream_t s, const EngineCfg& cfg, const Collect& col,
                            const int64_t* d_target, ChainState* st, float* vec, double* wel,
                            float* zreq, float* gout, double* peout, int64_t* gevals) override {
    mark_begin(s);
    const hipError_t e = launch(shape_code(waves(), rw), s, cfg, col, d_target, st, vec, wel,
                                zreq, gout, peout, gevals);
    mark_end(s);
    return e;
  }
};

Potential* make_hier(const float* x, const float* y, const int64_t* d_row_off,
                     const int64_t* d_elt_off, const int* d_dims, int groups, int K, int C,
                     int Dp, hipStream_t s, std::string* err) {
  if (K != 1 && K != 2 && K != 4 && K != 8 && K != 16) {
    *err = "hierarchical model: chains_per_group must be 1, 2, 4, 8 or 16";
    return nullptr;
  }
  std::vector<int> hd_dims(groups);
  std::vector<int64_t> hrow(groups + 1), helt(groups);
  if (hipMemcpy(hd_dims.data(), d_dims, sizeof(int) * groups, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hrow.data(), d_row_off, sizeof(int64_t) * (groups + 1), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(helt.data(), d_elt_off, sizeof(int64_t) * groups, hipMemcpyDeviceToHost) != hipSuccess) {
    *err = "hierarchical model: cannot read the group tables";
    return nullptr;
  }
  auto* p = new HierPotential();
  p->groups = groups;
  p->C = C;
  double work = 0.0;
  std::vector<double> w(groups);
  for (int i = 0; i < groups; ++i) {
    const int64_t n = hrow[i + 1] - hrow[i];
    if (hd_dims[i] < 1 || hd_dims[i] > kHierMaxD || n < 1) {
      *err = "hierarchical model: every group needs 1 <= D_i <= 64 and N_i >= 1";
      delete p;
      return nullptr;
    }
    p->maxd = std::max(p->maxd, hd_dims[i]);
    w[i] = (double)n * hd_dims[i];
    work += w[i];
  }
  p->flops = 4.0 * K * work;         // x.beta and X^T r per chain
  p->bytes = 4.0 * work + 4.0 * (double)(hrow[groups] - hrow[0]);  // X once per model + labels
  std::vector<int> ord(groups);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return w[a] > w[b]; });
  hipError_t e = hipMalloc(&p->order, sizeof(int) * groups);
  if (e == hipSuccess) e = hipMalloc(&p->dims, sizeof(int) * groups);
  if (e == hipSuccess) e = hipMalloc(&p->cdims, sizeof(int) * C);
  if (e == hipSuccess) e = hipMalloc(&p->row_off, sizeof(int64_t) * (groups + 1));
  if (e == hipSuccess) e = hipMalloc(&p->elt_off, sizeof(int64_t) * groups);
  if (e == hipSuccess) e = hipMemcpy(p->order, ord.data(), sizeof(int) * groups, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->dims, hd_dims.data(), sizeof(int) * groups, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->row_off, hrow.data(), sizeof(int64_t) * (groups + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->elt_off, helt.data(), sizeof(int64_t) * groups, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    *err = std::string("hierarchical model: ") + hipGetErrorString(e);
    delete p;
    return nullptr;
  }
  hipLaunchKernelGGL(hier_chain_dims, dim3((C + 255) / 256), dim3(256), 0, s, p->dims, K, C, p->cdims);
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    *err = std::string("hierarchical model: ") + hipGetErrorString(e);
    delete p;
    return nullptr;
  }
  p->hd = HierData{x, y, p->row_off, p->elt_off, p->dims, p->order, K, Dp, 0};
  return p;
}

}  // namespace nax
