// Bayesian logistic regression with one wavefront per chain, for few chains
// on small designs (the reference's CPU-sized configuration): the whole NUTS
// loop of a chain runs in one persistent launch, the chain's own wave
// evaluating U and dU (examples/covtype.py:66-71: coefs ~ Normal(0, s),
// obs ~ Bernoulli(logits = X coefs); BernoulliLogits discrete.py:244-263,
// binary_cross_entropy_with_logits distributions/util.py:317-320).  With a
// handful of chains the batched tile kernel cannot fill the GPU and every
// tick pays a launch round trip; here a leaf costs one wave's pass over X
// (L2-resident) and no launch.
//
// Lane = row: each lane takes rows lane, lane + 64, ... and keeps the
// chain's position (d <= DMAX) and its own partial X^T r in registers; one
// wave reduction per coordinate finishes the gradient.  The model reads a
// column-major copy of X made at set-up, so that the lanes' reads of one
// coordinate of consecutive rows are contiguous.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "nax_device.h"
#include "nax_engine.h"
#include "nax_internal.h"

namespace nax {

template <int DMAX>
struct LogisticChainModel {
  const float* xt;  // [d, n] column-major copy of X, fp32
  const float* y;   // [n]
  int64_t n;
  int d;
  float inv_s2;
  double log_norm;  // d * log(sqrt(2 pi) s)

  template <class G>
  __device__ double value_and_grad(const G& grp, const float* z, float* g) const {
    const int lane = grp.lane & (NAX_WAVE - 1);
    float zr[DMAX], gp[DMAX];
#pragma unroll
    for (int k = 0; k < DMAX; ++k) {
      zr[k] = k < d ? z[k] : 0.0f;
      gp[k] = 0.0f;
    }
    double u = 0.0;
    for (int64_t row = grp.lane; row < n; row += G::NT) {
      float xv[DMAX];
      float l0 = 0.0f, l1 = 0.0f;
#pragma unroll
      for (int k = 0; k < DMAX; ++k) {
        xv[k] = k < d ? xt[(int64_t)k * n + row] : 0.0f;
        if (k & 1) l1 = fmaf(xv[k], zr[k], l1);
        else l0 = fmaf(xv[k], zr[k], l0);
      }
      float r, uu;
      logistic_terms_f(l0 + l1, y[row], r, uu);
      u += (double)uu;
#pragma unroll
      for (int k = 0; k < DMAX; ++k) gp[k] = fmaf(xv[k], r, gp[k]);
    }
    double pr = 0.0;
#pragma unroll
    for (int k = 0; k < DMAX; ++k)
      if (k < d) pr += 0.5 * (double)zr[k] * zr[k] * inv_s2;
    // Transpose-reduce of the DMAX partial gradients over the 64 lanes: the
    // step with shuffle distance `off` keeps, in each lane, the half of its
    // coordinates selected by lane bit `off` and adds the partner lane's
    // copy of that half.  After log2(DMAX) steps lane l holds coordinate
    // k = l >> (6 - log2 DMAX), summed over the lanes that agree with it on
    // the bits not yet used; the remaining xor steps finish the sum.
    // 31 (DMAX 32) or 63 (DMAX 64) shuffles instead of 6 per coordinate.
    int n = DMAX, kbit = 0;
#pragma unroll
    for (int off = NAX_WAVE / 2; n > 1; off >>= 1, n >>= 1) {
      const int half = n / 2;
      const bool upper = (lane & off) != 0;
#pragma unroll
      for (int i = 0; i < half; ++i) {
        const float keep = upper ? gp[half + i] : gp[i];
        const float give = upper ? gp[i] : gp[half + i];
        gp[i] = keep + __shfl_xor(give, off, NAX_WAVE);
      }
      kbit = off;
    }
#pragma unroll
    for (int off = kbit >> 1; off > 0; off >>= 1) gp[0] += __shfl_xor(gp[0], off, NAX_WAVE);
    constexpr int shift = DMAX == 64 ? 0 : (DMAX == 32 ? 1 : -1);
    static_assert(shift >= 0, "DMAX is 32 or 64");
    const int k = lane >> shift;
    if constexpr (G::WPC == 1) {
      if ((lane & ((1 << shift) - 1)) == 0 && k < d) g[k] = gp[0] + z[k] * inv_s2;
    } else {
      // the waves' partial gradients meet in LDS, summed in wave order
      const int w = grp.lane / NAX_WAVE;
      __syncthreads();
      if ((lane & ((1 << shift) - 1)) == 0 && k < d) grp.redf[w * NAX_WAVE + k] = gp[0];
      __syncthreads();
      if (grp.lane < d) {
        float t = 0.0f;
#pragma unroll
        for (int i = 0; i < G::WPC; ++i) t += grp.redf[i * NAX_WAVE + grp.lane];
        g[grp.lane] = t + z[grp.lane] * inv_s2;
      }
    }
    return chain_sum(grp, u) + pr + log_norm;
  }
};

template <int DMAX>
struct LogisticChainPotential : Potential {
  LogisticChainModel<DMAX> m;
  float* xt = nullptr;  // owned column-major copy of X
  ~LogisticChainPotential() override { (void)hipFree(xt); }
  double flops = 0.0, bytes = 0.0;
  int dim() const override { return m.d; }
  const char* name() const override { return "logistic_chain"; }
  bool persistent() const override { return true; }
  double alg_flops(int C) const override { return 4.0 * (double)m.n * m.d * C; }
  double alg_bytes(int C) const override { return 4.0 * (double)m.n * (m.d + 1); }

  hipError_t eval(hipStream_t s, const float* z, int ldz, int C, double* pe, float* g,
                  int ldg) override {
    // one wave per chain unless a launch shape is fixed
    mark_begin(s);
    const hipError_t e = launch_group_eval(forced ? wpc : 1, s, m, z, ldz, C, pe, g, ldg);
    mark_end(s);
    return e;
  }
  hipError_t prepare_persistent(hipStream_t s, const EngineCfg& cfg, ChainState* st, float* vec,
                                double* wel, float* zreq, float* gout, double* peout) override {
    // one chain per workgroup; the waves per chain (rows of X split over
    // 64 wpc threads; the chain's vectors fit one wave, so with more than one
    // wave the first runs the state machine alone) are picked on the first
    // launch by timing a few iterations of each candidate from a scratch
    // copy of the state
    if (!forced && (wpc == 0 || tuned_chains != cfg.C)) {
      const std::string key = "logistic_chain n=" + std::to_string(m.n) + " d=" +
                              std::to_string(m.d) + " C=" + std::to_string(cfg.C) +
                              " dense=" + std::to_string(cfg.dense);
      tuned_chains = cfg.C;
      if (tuned_shape(key, &wpc, false)) {
        tuning = "logistic_chain: " + std::to_string(wpc) + " waves per chain (tuned earlier "
                 "for " + key + ")";
        return hipSuccess;
      }
      std::string log;
      wpc = tune_chain_groups<LogisticChainModel<DMAX>, true, true>(s, m, cfg, st, vec, wel,
                                                                     zreq, gout, peout, 4, &log);
      tuned_shape(key, &wpc, true);
      tuning = "logistic_chain: " + std::to_string(wpc) + " waves per chain (ms for 4 "
               "iterations:" + log + ")";
    }
    return hipSuccess;
  }
  hipError_t run_persistent(hipStream_t s, const EngineCfg& cfg, const Collect& col,
                            const int64_t* d_target, ChainState* st, float* vec, double* wel,
                            float* zreq, float* gout, double* peout, int64_t* gevals) override {
    mark_begin(s);
    hipError_t e = launch_chain_groups<LogisticChainModel<DMAX>, true, true>(
        wpc > 0 ? wpc : 1, s, m, cfg, col, d_target, st, vec, wel, zreq, gout, peout, gevals);
    mark_end(s);
    return e;
  }
  bool force_launch_shape(int shape) override {
    for (int w : kGroupSizes)
      if (w == shape) {
        wpc = shape;
        forced = true;
        tuning = "logistic_chain: " + std::to_string(wpc) + " waves per chain (fixed)";
        return true;
      }
    return false;
  }
  int wpc = 0, tuned_chains = 0;
  bool forced = false;
};

// xt[k n + i] = x[i d + k]
__global__ void transpose_rows_kernel(const float* __restrict__ x, int64_t n, int d,
                                      float* __restrict__ xt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * d) return;
  xt[(i % d) * n + i / d] = x[i];
}

// d <= 64; y is the caller's device array (kept, not copied), X is copied
// column-major.
Potential* make_logistic_chain(const float* x, const float* y, int64_t n, int d,
                               double prior_scale, hipStream_t s, std::string* err) {
  if (d < 1 || d > 64) {
    *err = "the per-chain logistic kernel supports 1 <= dim <= 64";
    return nullptr;
  }
  auto fill = [&](auto* p) -> Potential* {
    const int64_t nd = n * d;
    hipError_t e = hipMalloc(&p->xt, sizeof(float) * (size_t)(nd > 0 ? nd : 1));
    if (e == hipSuccess && nd > 0) {
      hipLaunchKernelGGL(transpose_rows_kernel, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0,
                         s, x, n, d, p->xt);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      *err = std::string("per-chain logistic model: ") + hipGetErrorString(e);
      delete p;
      return nullptr;
    }
    p->m.xt = p->xt;
    p->m.y = y;
    p->m.n = n;
    p->m.d = d;
    p->m.inv_s2 = (float)(1.0 / (prior_scale * prior_scale));
    p->m.log_norm = d * std::log(std::sqrt(2.0 * M_PI) * prior_scale);
    return p;
  };
  if (d <= 32) return fill(new LogisticChainPotential<32>());
  return fill(new LogisticChainPotential<64>());
}

}  // namespace nax
