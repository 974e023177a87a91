// Stochastic-volatility potential and its hand-written adjoint, one wavefront
// per chain (examples/stochastic_volatility.py:58-66, whose rates are
// a = 50, b = 0.1):
//   sigma ~ Exponential(a), nu ~ Exponential(b),
//   s ~ GaussianRandomWalk(sigma, T), r_t ~ StudentT(nu, 0, exp(s_t)).
// Log-densities: Exponential continuous.py:715-726, GaussianRandomWalk
// continuous.py:1104-1112, StudentT continuous.py:3229-3238; positive sites
// are unconstrained by ExpTransform (transforms.py:606-652, log|J| = x) as in
// potential_energy (numpyro/infer/util.py:300-358).
//
// Centred form (the reference's model), z = [log nu, s_0..s_{T-1}, log sigma]
// (sorted site names): the GRW term is a difference + sum, so its adjoint is a
// 3-point stencil and every t is independent -- lanes take t = lane + 64 k.
//
// Non-centred form, z = [log nu, log sigma, zeta_0..zeta_{T-1}] with
// s = sigma * cumsum(zeta): the adjoint of the prefix sum is a reverse scan,
// dU/dzeta_t = zeta_t + sigma * sum_{tau >= t} dU/ds_tau.  Lane l owns the
// contiguous segment [lL, (l+1)L), L = ceil(T/64) (about sqrt(T) for T up to
// ~4096): pass 1 sums each segment; a wave scan turns the sums into one
// checkpoint per segment (the prefix at its start); pass 2 walks the segment
// from its checkpoint computing the StudentT terms and dU/ds, whose segment
// totals a reverse wave scan turns into suffix checkpoints; pass 3 walks the
// segment again, recomputing s_t and dU/ds_t from the checkpoints (nothing
// per-t is stored) and forms each suffix as (segment total - running prefix).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "nax_device.h"
// T + 2 latent entries per chain: ~47 per lane at T = 3000
#define NAX_FOR_D_UNROLL 8
#include "nax_engine.h"
#include "nax_internal.h"

namespace nax {

// digamma for x > 0: recurrence psi(x) = psi(x + 1) - 1/x up to x >= 6, then
// the asymptotic series ln x - 1/(2x) - sum_k B_2k / (2k x^2k)
// (Abramowitz & Stegun 6.3.5 and 6.3.18).
__device__ __forceinline__ double digamma_pos(double x) {
  double acc = 0.0;
  while (x < 6.0) {
    acc -= 1.0 / x;
    x += 1.0;
  }
  const double i2 = 1.0 / (x * x);
  const double series =
      i2 * (1.0 / 12 - i2 * (1.0 / 120 - i2 * (1.0 / 252 - i2 * (1.0 / 240 - i2 * (1.0 / 132)))));
  return acc + log(x) - 0.5 / x - series;
}

constexpr double kLogPi = 1.1447298858494002;
// Longest non-centred segment per thread kept in registers across the passes
// (T = 3000 on 4 waves per chain: 12 steps per thread).
constexpr int kSegRegs = 16;
constexpr double kLog2Pi = 1.8378770664093453;

struct SVModel {
  const float* r;  // returns [T]
  int T;
  int noncentred;
  double sigma_rate;  // sigma ~ Exponential(sigma_rate)
  double nu_rate;     // nu ~ Exponential(nu_rate)
  int softplus;       // positive sites by SoftplusTransform instead of ExpTransform

  // T-dependent constant part of U and of dU/dnu for the StudentT block.
  // The two lgamma and two digamma values are evaluated on lanes 0 and 1 at
  // once (same code, argument by lane), then read out of those lanes.
  __device__ void studentt_consts(double nu, double& u_const, double& dnu_const) const {
    const int ln = (int)(threadIdx.x & (NAX_WAVE - 1));
    const double x = (ln & 1) ? 0.5 * (nu + 1.0) : 0.5 * nu;
    const double lg = lgamma(x), dg = digamma_pos(x);
    const double lg0 = read_lane(lg, 0), lg1 = read_lane(lg, 1);
    const double dg0 = read_lane(dg, 0), dg1 = read_lane(dg, 1);
    u_const = T * (0.5 * log(nu) + 0.5 * kLogPi + lg0 - lg1);
    dnu_const = T * (0.5 / nu + 0.5 * dg0 - 0.5 * dg1);
  }

  template <class G>
  __device__ __forceinline__ double value_and_grad(const G& grp, const float* z, float* g) const {
    return noncentred ? nc(grp, z, g) : centred(grp, z, g);
  }

  template <class G>
  __device__ __forceinline__ double centred(const G& grp, const float* z, float* g) const {
    const int lane = grp.lane;
    const double u_nu = z[0], u_sig = z[T + 1];
    const PosTf tn = positive_tf(u_nu, softplus), ts = positive_tf(u_sig, softplus);
    const double nu = tn.x, sig = ts.x;
    const float nuf = (float)nu, nu1 = (float)(nu + 1.0);
    const float inv_s2 = (float)(1.0 / (sig * sig));
    const float inv_nu = 1.0f / nuf;
    // dU/dnu's data term sum_t 0.5 a_t - 0.5 (nu+1) b_t / nu: sum a and b per
    // lane, scale once (nu is constant within an evaluation)
    double U = 0.0, S2 = 0.0, asum = 0.0, bsum = 0.0;
#pragma unroll 4
    for (int t = lane; t < T; t += G::NT) {
      const float s = z[1 + t];
      const float dprev = s - (t > 0 ? z[t] : 0.0f);
      const float dnext = (t + 1 < T) ? (z[2 + t] - s) : 0.0f;
      const float rr = r[t];
      const float y2 = rr * rr * __expf(-2.0f * s);
      const float a = log1pf(y2 * inv_nu);
      const float b = y2 / (nuf + y2);
      U += (double)(0.5f * dprev * dprev * inv_s2) + 0.5 * (double)nu1 * (double)a + (double)s;
      S2 += (double)dprev * (double)dprev;
      asum += (double)a;
      bsum += (double)b;
      g[1 + t] = (dprev - dnext) * inv_s2 + 1.0f - nu1 * b;
    }
    double v[4] = {U, S2, asum, bsum};
    chain_sum_n(grp, v);
    U = v[0];
    S2 = v[1];
    const double dnu = 0.5 * v[2] - 0.5 * (double)nu1 * v[3] / nu;
    double uc, dnc;
    studentt_consts(nu, uc, dnc);
    U += uc + T * ((softplus ? log(sig) : u_sig) + 0.5 * kLog2Pi)  // GRW normaliser
         - log(sigma_rate) + sigma_rate * sig - ts.ladj     // Exponential(a) + log|J|
         - log(nu_rate) + nu_rate * nu - tn.ladj;           // Exponential(b) + log|J|
    if (lane == 0) {
      // dU/du = dU/dx dx/du - d log|J| / du for u = log nu, log sigma (or
      // softplus^-1); for ExpTransform sig (-S2 / sig^3 + T / sig + a) - 1
      g[0] = (float)(tn.dx * (dnu + dnc + nu_rate) - tn.dladj);
      g[T + 1] = softplus
          ? (float)(ts.dx * (-S2 / (sig * sig * sig) + T / sig + sigma_rate) - ts.dladj)
          : (float)(-S2 / (sig * sig) + T + sigma_rate * sig - 1.0);
    }
    return U;
  }

  template <class G>
  __device__ __forceinline__ double nc(const G& grp, const float* z, float* g) const {
    const int lane = grp.lane;
    const double u_nu = z[0], u_sig = z[1];
    const PosTf tn = positive_tf(u_nu, softplus), ts = positive_tf(u_sig, softplus);
    const double nu = tn.x, sig = ts.x;
    const float nuf = (float)nu, nu1 = (float)(nu + 1.0);
    const float* zeta = z + 2;
    float* gz = g + 2;
    const int L = (T + G::NT - 1) / G::NT;
    const int t0 = lane * L;
    const int t1 = min(T, t0 + L);
    if (L <= kSegRegs) return nc_short(grp, z, g, tn, ts, t0, t1);
    // pass 1: segment sums -> prefix checkpoints
    double seg = 0.0;
#pragma unroll 8
    for (int t = t0; t < t1; ++t) seg += (double)zeta[t];
    const double cp = chain_inclusive_scan(grp, seg) - seg;
    // pass 2: StudentT terms and dU/ds_t = 1 - (nu+1) y^2 / (nu + y^2)
    const float inv_nu = 1.0f / nuf;
    double U = 0.0, asum = 0.0, bsum = 0.0, A = 0.0, Asc = 0.0, Z2 = 0.0;
    double c = cp;
#pragma unroll 4
    for (int t = t0; t < t1; ++t) {
      const float zt = zeta[t];
      c += (double)zt;
      const float s = (float)(sig * c);
      const float rr = r[t];
      const float y2 = rr * rr * __expf(-2.0f * s);
      const float a = log1pf(y2 * inv_nu);
      const float b = y2 / (nuf + y2);
      const float ds = 1.0f - nu1 * b;
      U += 0.5 * (double)nu1 * (double)a + (double)s;
      asum += (double)a;
      bsum += (double)b;
      A += (double)ds;
      Asc += (double)ds * c;
      Z2 += (double)zt * (double)zt;
    }
    const double Atot = chain_sum(grp, A);
    const double after = Atot - chain_inclusive_scan(grp, A);  // sum over later segments
    // pass 3: dU/dzeta_t = zeta_t + sigma * (suffix of dU/ds from t)
    double P = 0.0;
    c = cp;
#pragma unroll 4
    for (int t = t0; t < t1; ++t) {
      const float zt = zeta[t];
      c += (double)zt;
      const float s = (float)(sig * c);
      const float rr = r[t];
      const float y2 = rr * rr * __expf(-2.0f * s);
      const float ds = 1.0f - nu1 * (y2 / (nuf + y2));
      gz[t] = (float)((double)zt + sig * (after + A - P));
      P += (double)ds;
    }
    return nc_tail(grp, g, tn, ts, U, Z2, asum, bsum, Asc);
  }

  // The group sums and the scalar terms of the non-centred form.
  template <class G>
  __device__ __forceinline__ double nc_tail(const G& grp, float* g, const PosTf& tn,
                                            const PosTf& ts, double U, double Z2, double asum,
                                            double bsum, double Asc) const {
    const double nu = tn.x, sig = ts.x;
    const float nu1 = (float)(nu + 1.0);
    double v[5] = {U, Z2, asum, bsum, Asc};
    chain_sum_n(grp, v);
    U = v[0] + 0.5 * v[1];
    const double dnu = 0.5 * v[2] - 0.5 * (double)nu1 * v[3] / nu;
    const double asc = v[4];
    double uc, dnc;
    studentt_consts(nu, uc, dnc);
    U += uc + T * 0.5 * kLog2Pi                         // N(0,1) normaliser of zeta
         - log(sigma_rate) + sigma_rate * sig - ts.ladj
         - log(nu_rate) + nu_rate * nu - tn.ladj;
    if (grp.lane == 0) {
      g[0] = (float)(tn.dx * (dnu + dnc + nu_rate) - tn.dladj);
      g[1] = softplus ? (float)(ts.dx * (asc + sigma_rate) - ts.dladj)
                      : (float)(sig * asc + sigma_rate * sig - 1.0);
    }
    return U;
  }

  // The non-centred form for segments of at most kSegRegs steps: the three
  // passes of nc() with the segment's zeta, returns and dU/ds kept in
  // registers between them (the same expressions in the same order).
  template <class G>
  __device__ __forceinline__ double nc_short(const G& grp, const float* z, float* g,
                                             const PosTf& tn, const PosTf& ts, int t0,
                                             int t1) const {
    const double nu = tn.x, sig = ts.x;
    const float nuf = (float)nu, nu1 = (float)(nu + 1.0);
    const float* zeta = z + 2;
    float* gz = g + 2;
    float zk[kSegRegs], rk[kSegRegs], dk[kSegRegs];
#pragma unroll
    for (int k = 0; k < kSegRegs; ++k) {
      const int t = t0 + k;
      zk[k] = t < t1 ? zeta[t] : 0.0f;
      rk[k] = t < t1 ? r[t] : 0.0f;
    }
    double seg = 0.0;
#pragma unroll
    for (int k = 0; k < kSegRegs; ++k)
      if (t0 + k < t1) seg += (double)zk[k];
    const double cp = chain_inclusive_scan(grp, seg) - seg;
    const float inv_nu = 1.0f / nuf;
    double U = 0.0, asum = 0.0, bsum = 0.0, A = 0.0, Asc = 0.0, Z2 = 0.0;
    double c = cp;
#pragma unroll
    for (int k = 0; k < kSegRegs; ++k) {
      dk[k] = 0.0f;
      if (t0 + k < t1) {
        const float zt = zk[k];
        c += (double)zt;
        const float s = (float)(sig * c);
        const float rr = rk[k];
        const float y2 = rr * rr * __expf(-2.0f * s);
        const float a = log1pf(y2 * inv_nu);
        const float b = y2 / (nuf + y2);
        const float ds = 1.0f - nu1 * b;
        dk[k] = ds;
        U += 0.5 * (double)nu1 * (double)a + (double)s;
        asum += (double)a;
        bsum += (double)b;
        A += (double)ds;
        Asc += (double)ds * c;
        Z2 += (double)zt * (double)zt;
      }
    }
    const double Atot = chain_sum(grp, A);
    const double after = Atot - chain_inclusive_scan(grp, A);  // sum over later segments
    double P = 0.0;
#pragma unroll
    for (int k = 0; k < kSegRegs; ++k) {
      const int t = t0 + k;
      if (t < t1) {
        gz[t] = (float)((double)zk[k] + sig * (after + A - P));
        P += (double)dk[k];
      }
    }
    return nc_tail(grp, g, tn, ts, U, Z2, asum, bsum, Asc);
  }
};

struct SVPotential : Potential {
  SVModel m;
  float* r = nullptr;  // owned copy of the returns
  ~SVPotential() override { hipFree(r); }
  int dim() const override { return m.T + 2; }
  const char* name() const override { return "sv"; }
  bool set_positive_transform(int kind) override {
    m.softplus = kind == 1;
    return true;
  }
  bool persistent() const override { return true; }
  double alg_bytes(int C) const override { return (double)C * 4.0 * (2.0 * (m.T + 2) + m.T); }

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
    // one chain per workgroup, its vector slots in LDS; the waves per chain
    // (a chain's time steps and vector passes split over 64 wpc threads)
    // are picked on the first launch by timing a few iterations of each
    // candidate from a scratch copy of the state
    if (!forced && (wpc == 0 || tuned_chains != cfg.C)) {
      const std::string key = "sv T=" + std::to_string(m.T) + " nc=" +
                              std::to_string(m.noncentred) + " C=" + std::to_string(cfg.C) +
                              " dense=" + std::to_string(cfg.dense);
      tuned_chains = cfg.C;
      if (tuned_shape(key, &wpc, false)) {
        tuning = "sv: " + std::to_string(wpc) + " waves per chain (tuned earlier for " + key +
                 ")";
        return hipSuccess;
      }
      std::string log;
      wpc = tune_chain_groups<SVModel, true>(s, m, cfg, st, vec, wel, zreq, gout, peout, 2, &log);
      tuned_shape(key, &wpc, true);
      tuning = "sv: " + std::to_string(wpc) + " waves per chain (ms for 2 iterations:" + log +
               ")";
    }
    return hipSuccess;
  }
  hipError_t run_persistent(hipStream_t s, const EngineCfg& cfg, const Collect& col,
                            const int64_t* d_target, ChainState* st, float* vec, double* wel,
                            float* zreq, float* gout, double* peout, int64_t* gevals) override {
    mark_begin(s);
    hipError_t e = launch_chain_groups<SVModel, true>(wpc > 0 ? wpc : 1, s, m, cfg, col, d_target,
                                                     st, vec, wel, zreq, gout, peout, gevals);
    mark_end(s);
    return e;
  }
  bool force_launch_shape(int shape) override {
    for (int w : kGroupSizes)
      if (w == shape) {
        wpc = shape;
        forced = true;
        tuning = "sv: " + std::to_string(wpc) + " waves per chain (fixed)";
        return true;
      }
    return false;
  }
  int wpc = 0, tuned_chains = 0;
  bool forced = false;
};

Potential* make_sv(const float* returns, int T, int noncentred, double sigma_rate, double nu_rate,
                   hipStream_t s, std::string* err) {
  if (T < 2) {
    *err = "the stochastic-volatility model needs T >= 2";
    return nullptr;
  }
  if (!(sigma_rate > 0.0) || !(nu_rate > 0.0) || !std::isfinite(sigma_rate) ||
      !std::isfinite(nu_rate)) {
    *err = "the stochastic-volatility model needs finite positive Exponential rates";
    return nullptr;
  }
  auto* p = new SVPotential();
  hipError_t e = hipMalloc(&p->r, sizeof(float) * T);
  if (e == hipSuccess)
    e = hipMemcpyAsync(p->r, returns, sizeof(float) * T, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    *err = std::string("sv model: ") + hipGetErrorString(e);
    delete p;
    return nullptr;
  }
  p->m.r = p->r;
  p->m.T = T;
  p->m.noncentred = noncentred;
  p->m.sigma_rate = sigma_rate;
  p->m.nu_rate = nu_rate;
  p->m.softplus = 0;
  return p;
}

}  // namespace nax
