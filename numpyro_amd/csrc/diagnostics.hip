// On-device MCMC diagnostics over collected draws x [C chains][S draws][P]
// (the layout nax_run collects z in): effective sample size with Geyer's
// initial monotone sequence and split Gelman-Rubin, per parameter
// (numpyro/diagnostics.py:158-203, 64-80, 29-42).  The reference computes the
// autocovariance by FFT; here one workgroup per parameter stages each chain's
// draws in LDS as fp64 and sums lag products directly (S^2 / 2 products per
// chain: S <= 4096), which gives the same (biased) autocovariance
// sum_t xc_t xc_{t+k} / S.  Chain means and variances are combined with
// Welford updates across chains, so no per-chain arrays are kept.
#include <hip/hip_runtime.h>

#include <string>

#include "nax_device.h"
#include "nax_internal.h"

namespace nax {

constexpr int kDiagMaxDraws = 4096;
constexpr int kDiagThreads = 256;

__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < kDiagThreads / NAX_WAVE; ++i) t += red[i];
  return t;
}

// Welford over chains of (mean, variance) pairs.
struct ChainAgg {
  int n = 0;
  double mean_of_means = 0.0, m2_of_means = 0.0, sum_var = 0.0;
  __device__ void add(double m, double v) {
    n += 1;
    const double d = m - mean_of_means;
    mean_of_means += d / n;
    m2_of_means += d * (m - mean_of_means);
    sum_var += v;
  }
  // _compute_chain_variance_stats: (var_within, var_estimator) for chains of
  // length len
  __device__ void stats(int len, double& vw, double& ve) const {
    vw = sum_var / n;
    ve = vw * (len - 1) / len;
    if (n > 1) ve += m2_of_means / (n - 1);
    else vw = ve;
  }
};

__global__ void __launch_bounds__(kDiagThreads)
diagnostics_kernel(const float* __restrict__ x, int C, int S, int P, double* __restrict__ ess,
                   double* __restrict__ rhat) {
  __shared__ double xs[kDiagMaxDraws];
  __shared__ double gsum[kDiagMaxDraws];
  __shared__ double red[kDiagThreads / NAX_WAVE];
  const int p = blockIdx.x;
  const int tid = threadIdx.x;
  for (int k = tid; k < S; k += kDiagThreads) gsum[k] = 0.0;
  ChainAgg full, split;  // identical on every thread
  const int h = S / 2;
  for (int c = 0; c < C; ++c) {
    __syncthreads();  // previous chain's lag sums are done with xs
    double s = 0.0;
    for (int t = tid; t < S; t += kDiagThreads) {
      const double v = x[((size_t)c * S + t) * P + p];
      xs[t] = v;
      s += v;
    }
    const double mean = block_sum(s, red) / S;
    double ss = 0.0, s1 = 0.0, s2 = 0.0;
    for (int t = tid; t < S; t += kDiagThreads) {
      const double xc = xs[t] - mean;
      xs[t] = xc;
      ss += xc * xc;
      if (t < h) s1 += xc;
      if (t >= S - h) s2 += xc;
    }
    ss = block_sum(ss, red);
    s1 = block_sum(s1, red);
    s2 = block_sum(s2, red);
    full.add(mean, ss / (S - 1));
    if (h >= 2) {
      // halves [0, h) and [S - h, S): centred on their own means
      const double m1 = s1 / h, m2 = s2 / h;
      double q1 = 0.0, q2 = 0.0;
      for (int t = tid; t < h; t += kDiagThreads) {
        const double a = xs[t] - m1, b = xs[S - h + t] - m2;
        q1 += a * a;
        q2 += b * b;
      }
      q1 = block_sum(q1, red);
      q2 = block_sum(q2, red);
      split.add(mean + m1, q1 / (h - 1));
      split.add(mean + m2, q2 / (h - 1));
    }
    // biased autocovariance at lags k: sum_t xc_t xc_{t+k} / S
    for (int k = tid; k < S; k += kDiagThreads) {
      double a = 0.0;
      for (int t = 0; t + k < S; ++t) a += xs[t] * xs[t + k];
      gsum[k] += a / S;
    }
  }
  __syncthreads();
  if (tid == 0) {
    double vw, ve;
    full.stats(S, vw, ve);
    // rho_k = 1 - (var_within - mean_c gamma_ck) / var_estimator, rho_0 = 1;
    // Rho_k = rho_2k + rho_2k+1; initial monotone sequence on Rho_1..
    double tau = 0.0, run = 0.0;
    const int npair = S / 2;
    for (int k = 0; k < npair; ++k) {
      const double r0 = (2 * k == 0) ? 1.0 : 1.0 - (vw - gsum[2 * k] / C) / ve;
      const double r1 = 1.0 - (vw - gsum[2 * k + 1] / C) / ve;
      double R = r0 + r1;
      if (k >= 1) {
        R = R > 0.0 ? R : 0.0;
        run = (k == 1) ? R : (R < run ? R : run);
        R = run;
      }
      tau += R;
    }
    tau = -1.0 + 2.0 * tau;
    ess[p] = (double)C * S / tau;
    if (rhat) {
      double svw, sve;
      if (h >= 2) {
        split.stats(h, svw, sve);
        rhat[p] = sqrt(sve / svw);
      } else {
        rhat[p] = NAN;
      }
    }
  }
}

}  // namespace nax

extern "C" int nax_diagnostics(const float* d_x, int32_t chains, int32_t draws, int32_t params,
                               double* d_ess, double* d_split_rhat) {
  using namespace nax;
  if (!d_x || !d_ess || chains < 1 || draws < 2 || params < 1) {
    set_error("nax_diagnostics: need x, ess, chains >= 1, draws >= 2, params >= 1");
    return NAX_EINVAL;
  }
  if (draws > kDiagMaxDraws) {
    set_error("nax_diagnostics: at most 4096 draws per chain");
    return NAX_EINVAL;
  }
  hipLaunchKernelGGL(diagnostics_kernel, dim3(params), dim3(kDiagThreads), 0, 0, d_x, chains,
                     draws, params, d_ess, d_split_rhat);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    set_error(std::string("nax_diagnostics: ") + hipGetErrorString(e));
    return NAX_EHIP;
  }
  return NAX_OK;
}
