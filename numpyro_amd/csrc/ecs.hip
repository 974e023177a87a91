// HMCECS for logistic regression: the subsampled potential with the
// difference estimator (Taylor proxy of degree 1 or 2 around reference
// coefficients) or the plain n/m-scaled subsample log-likelihood, and the
// Gibbs block update of each chain's subsample indices.
//
//   estimator    numpyro/contrib/ecs_proxies.py:23-50
//   block update numpyro/contrib/ecs_proxies.py:58-92
//   Taylor proxy numpyro/contrib/ecs_proxies.py:95-250
//   Gibbs step   numpyro/infer/hmc_gibbs.py:640-681
//   model        examples/covtype.py:66-71 with subsample_size,
//                run under HMCECS in examples/covtype.py:154-165
//
// For the logistic likelihood the proxy of row i needs four numbers at the
// reference point: l_ref = x_i.beta_ref, ll_ref = log Bern(y_i | l_ref),
// r = y_i - sigmoid(l_ref) (gradient r x_i) and w = sigmoid(l_ref)(1 -
// sigmoid(l_ref)) (Hessian -w x_i x_i^T), so the per-row proxy is
// ll_ref + r dl - w dl^2 / 2 with dl = x_i.(beta - beta_ref): no per-row
// D-vector or DxD matrix is stored.  The full-data proxy uses LL = sum ll_ref,
// G = sum r x_i and H = -sum w x_i x_i^T, computed once at set-up.
//
// Layout: X row-major [n, d] fp32 (the caller's buffer), y [n] fp32, the
// per-row scalars as one 32-byte record per row (a gathered row costs one
// 32-byte read, not four cache lines), H [d, d] fp64.  Subsample indices [C, m]
// int32 and the HMCECS state keys [C, 2] are caller-owned device buffers.
#include <hip/hip_runtime.h>
#include <math.h>

#include <string>

#include "nax_device.h"
#include "nax_internal.h"

namespace nax {

constexpr int kEcsThreads = 256;
constexpr int kEcsMaxD = 512;
constexpr int kEcsMaxM = 4096;
constexpr int kEcsSplit = 256;  // row splits of the set-up reductions

// Per-row Taylor terms.  y (0 or 1) is carried in the sign bit of w (w >= 0),
// so the proxy's gathered reads of a row take one 32-byte record.
struct alignas(32) EcsRow {
  double l_ref, ll_ref, r, w_y;
};

__device__ __forceinline__ double ecs_w(const EcsRow& q) { return fabs(q.w_y); }
__device__ __forceinline__ double ecs_y(const EcsRow& q) { return signbit(q.w_y) ? 1.0 : 0.0; }

struct EcsArgs {
  const float* x;
  const float* y;
  int64_t n;
  int d, m, proxy;  // proxy: 0 none, 1 or 2 (Taylor degree)
  double prior_scale;
  const double* bref;
  const EcsRow* rows;  // [n]
  const double* G;   // [d]
  const double* H;   // [d, d]
  double LL;
};

__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int wv = threadIdx.x / NAX_WAVE, ln = threadIdx.x % NAX_WAVE;
  __syncthreads();
  if (ln == 0) red[wv] = v;
  __syncthreads();
  double t = 0;
  for (int i = 0; i < kEcsThreads / NAX_WAVE; ++i) t += red[i];
  return t;
}

// log Bernoulli(y | logits l) in fp64: y l - softplus(l).
__device__ __forceinline__ double bern_ll(double l, double y) {
  return y * l - (fmax(l, 0.0) + log1p(exp(-fabs(l))));
}

// log Bernoulli(y | logits l) and y - sigmoid(l) in fp64 from one
// e = exp(-|l|): softplus(l) = max(l, 0) + log1p(e), sigmoid(l) = 1 / (1 + e)
// for l >= 0 and e / (1 + e) otherwise.
__device__ __forceinline__ void bern_terms(double l, double y, double& ll, double& dy) {
  const double e = exp(-fabs(l));
  ll = y * l - (fmax(l, 0.0) + log1p(e));
  dy = y - (l >= 0.0 ? 1.0 : e) / (1.0 + e);
}

// ---- set-up at the reference point ------------------------------------------

// One wavefront per row: l_ref, ll_ref, r, w (with the sign of -y).
__global__ __launch_bounds__(kEcsThreads) void ecs_rows_kernel(EcsArgs a, EcsRow* rows) {
  const int ln = threadIdx.x % NAX_WAVE;
  const int64_t waves = (int64_t)gridDim.x * (kEcsThreads / NAX_WAVE);
  for (int64_t i = (int64_t)blockIdx.x * (kEcsThreads / NAX_WAVE) + threadIdx.x / NAX_WAVE;
       i < a.n; i += waves) {
    const float* xi = a.x + i * a.d;
    double s = 0;
    for (int k = ln; k < a.d; k += NAX_WAVE) s += (double)xi[k] * a.bref[k];
    s = wave_sum(s);
    if (ln == 0) {
      const double y = a.y[i];
      const double p = expit(s);
      const double w = a.proxy == 2 ? p * (1.0 - p) : 0.0;
      EcsRow q;
      q.l_ref = s;
      q.ll_ref = bern_ll(s, y);
      q.r = y - p;
      q.w_y = y > 0.5 ? -w : w;
      rows[i] = q;
    }
  }
}

// Partial sums over row split blockIdx.y: part[s][k] = sum r_i x_ik (k < d),
// part[s][d] = sum ll_ref_i.
__global__ __launch_bounds__(kEcsThreads) void ecs_grad_sum_kernel(EcsArgs a, double* part) {
  const int64_t rows = (a.n + kEcsSplit - 1) / kEcsSplit;
  const int64_t i0 = blockIdx.y * rows, i1 = min(a.n, i0 + rows);
  const int k = blockIdx.x * kEcsThreads + threadIdx.x;
  if (k > a.d) return;
  double s = 0;
  if (k < a.d)
    for (int64_t i = i0; i < i1; ++i) s += a.rows[i].r * (double)a.x[i * a.d + k];
  else
    for (int64_t i = i0; i < i1; ++i) s += a.rows[i].ll_ref;
  part[(int64_t)blockIdx.y * (a.d + 1) + k] = s;
}

// Partial Hessian over row split blockIdx.z for the 64x64 tile (blockIdx.x,
// blockIdx.y): part[s][i][j] = -sum w_r x_ri x_rj.  Rows are staged 32 at a
// time through LDS; each thread owns a 4x4 block of the tile.
__global__ __launch_bounds__(kEcsThreads) void ecs_hess_kernel(EcsArgs a, double* part) {
  __shared__ float xi[32][64];
  __shared__ float xj[32][64];
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  const int64_t rows = (a.n + kEcsSplit - 1) / kEcsSplit;
  const int64_t r0 = blockIdx.z * rows, r1 = min(a.n, r0 + rows);
  const int ti = threadIdx.x / 16, tj = threadIdx.x % 16;
  double acc[4][4] = {};
  for (int64_t rb = r0; rb < r1; rb += 32) {
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * 64; e += kEcsThreads) {
      const int rr = e / 64, cc = e % 64;
      const int64_t row = rb + rr;
      const bool ok = row < r1;
      const float wr = ok ? (float)ecs_w(a.rows[row]) : 0.0f;
      xi[rr][cc] = (ok && i0 + cc < a.d) ? wr * a.x[row * a.d + i0 + cc] : 0.0f;
      xj[rr][cc] = (ok && j0 + cc < a.d) ? a.x[row * a.d + j0 + cc] : 0.0f;
    }
    __syncthreads();
    for (int rr = 0; rr < 32; ++rr) {
      float p[4], q[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) { p[u] = xi[rr][ti * 4 + u]; q[u] = xj[rr][tj * 4 + u]; }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += (double)(p[u] * q[v]);
    }
  }
  double* out = part + (int64_t)blockIdx.z * a.d * a.d;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + ti * 4 + u, j = j0 + tj * 4 + v;
      if (i < a.d && j < a.d) out[(int64_t)i * a.d + j] = -acc[u][v];
    }
}

// out[k] = sum_s part[s][k], k < len.
__global__ void ecs_reduce_kernel(const double* part, int nsplit, int64_t len, double* out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= len) return;
  double s = 0;
  for (int i = 0; i < nsplit; ++i) s += part[(int64_t)i * len + k];
  out[k] = s;
}

// ---- the subsampled potential ------------------------------------------------

// One workgroup per chain (act[blockIdx.x] when a list is given), one pass
// over the subsample.  Each wavefront takes ROWS rows at a time, lanes
// over the coordinates (the row values stay in registers), reduces l = x.beta
// for each row, and lane r forms the fp64 per-row terms of row r: diff_r and
// c_r = d diff_r / d l_r.  The estimator's weight on diff_j,
// n/m (1 - n/m (diff_j - mean)), is n/m (1 + n/m mean) - (n/m)^2 diff_j, so
// the gradient needs only S1 = sum c_j x_j and S2 = sum diff_j c_j x_j, both
// accumulated from the registers while the mean is still unknown.
// NCH: 64-column chunks of a row held in registers (d <= 64 NCH); ROWS rows
// per wavefront step (fewer for wide rows, to stay within 128 VGPRs).
template <int NCH, int ROWS>
__global__ __launch_bounds__(kEcsThreads, NCH <= 4 ? 4 : 2) void ecs_eval_kernel(
    EcsArgs a, const int* __restrict__ idx_all, const float* __restrict__ z, int ldz,
    const int* act, const int* nact, double* pe, float* g, int ldg) {
  constexpr int NW = kEcsThreads / NAX_WAVE;
  __shared__ float beta[kEcsMaxD];
  __shared__ double dlt[kEcsMaxD];
  __shared__ double part[2][NW][kEcsMaxD];
  __shared__ double red[NW];
  __shared__ double shift[NW];
  int c = blockIdx.x;
  if (act) {
    if (c >= *nact) return;
    c = act[c];
  }
  const int d = a.d, m = a.m;
  const float* zc = z + (int64_t)c * ldz;
  const int* idx = idx_all + (int64_t)c * m;
  for (int k = threadIdx.x; k < kEcsMaxD; k += kEcsThreads) {
    beta[k] = k < d ? zc[k] : 0.0f;
    dlt[k] = (a.proxy && k < d) ? (double)zc[k] - a.bref[k] : 0.0;
  }
  __syncthreads();
  const int wv = threadIdx.x / NAX_WAVE, ln = threadIdx.x % NAX_WAVE;
  const int nch = (d + NAX_WAVE - 1) / NAX_WAVE;
  const double nm = (double)a.n / m;
  // S1 and S2 - d0 S1 in fp32, d0 = the diff of this wave's first row (so
  // the final combination has no large cancelling terms)
  float s1[NCH], s2[NCH];
#pragma unroll
  for (int t = 0; t < NCH; ++t) s1[t] = s2[t] = 0.0f;
  double sd = 0, sd2 = 0;  // lanes 0..ROWS-1: sum diff, sum diff^2
  double d0 = 0.0;
  auto load_rows = [&](int jb, float (&dst)[ROWS][NCH]) {
    const float* xr[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) xr[r] = a.x + (int64_t)idx[min(jb + r, m - 1)] * d;
#pragma unroll
    for (int t = 0; t < NCH; ++t) {
      const int k = ln + t * NAX_WAVE;
      const bool ok = t < nch && k < d;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) dst[r][t] = ok ? xr[r][k] : 0.0f;
    }
  };
  float xv[ROWS][NCH];
  int j0 = wv * ROWS;
  if (j0 < m) load_rows(j0, xv);
  for (bool first = true; j0 < m; j0 += NW * ROWS, first = false) {
    // the next step's rows are in flight while this step is reduced
    float xn[ROWS][NCH];
    const int jn = j0 + NW * ROWS;
    if (jn < m) load_rows(jn, xn);
    float s[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) s[r] = 0.0f;
#pragma unroll
    for (int t = 0; t < NCH; ++t) {
      const int k = ln + t * NAX_WAVE;
      const float b = (t < nch && k < d) ? beta[k] : 0.0f;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) s[r] = fmaf(xv[r][t], b, s[r]);
    }
    float mine = 0.0f;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const float t = wave_sum(s[r]);
      if (ln == r) mine = t;
    }
    const int j = j0 + ln;
    double df = 0, cc = 0;
    if (ln < ROWS && j < m) {
      const int64_t i = idx[j];
      const double l = mine;
      if (a.proxy) {
        const EcsRow q = a.rows[i];
        const double y = ecs_y(q), w = ecs_w(q);
        const double dl = l - q.l_ref;
        double ll, dy;
        bern_terms(l, y, ll, dy);
        df = ll - (q.ll_ref + q.r * dl - 0.5 * w * dl * dl);
        cc = dy - q.r + w * dl;
      } else {
        double dy;
        bern_terms(l, a.y[i], df, dy);
        cc = nm * dy;
      }
      sd += df;
      sd2 += df * df;
    }
    if (first) d0 = __shfl(df, 0, NAX_WAVE);
    const float ccf = (float)cc, dcf = (float)((df - d0) * cc);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const float cr = __shfl(ccf, r, NAX_WAVE);
      const float dcr = __shfl(dcf, r, NAX_WAVE);
#pragma unroll
      for (int t = 0; t < NCH; ++t) {
        s1[t] = fmaf(cr, xv[r][t], s1[t]);
        s2[t] = fmaf(dcr, xv[r][t], s2[t]);
      }
    }
    if (jn < m) {
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
#pragma unroll
        for (int t = 0; t < NCH; ++t) xv[r][t] = xn[r][t];
    }
  }
#pragma unroll
  for (int t = 0; t < NCH; ++t) {
    const int k = ln + t * NAX_WAVE;
    if (t < nch && k < d) {
      part[0][wv][k] = s1[t];
      part[1][wv][k] = s2[t];
    }
  }
  if (ln == 0) shift[wv] = d0;
  // (H delta)_k for this thread's coordinates k = tid + 256 q, kept for the
  // gradient; H is symmetric, so column reads are coalesced
  double hk[kEcsMaxD / kEcsThreads];
#pragma unroll
  for (int q = 0; q < kEcsMaxD / kEcsThreads; ++q) {
    hk[q] = 0.0;
    const int k = threadIdx.x + q * kEcsThreads;
    if (a.proxy && k < d)
      for (int l2 = 0; l2 < d; ++l2) hk[q] += a.H[(int64_t)l2 * d + k] * dlt[l2];
  }
  const double mean = block_sum(sd, red) / m;  // block_sum synchronises first
  double est, w1, w2;
  if (a.proxy) {
    const double var = fmax(block_sum(sd2, red) / m - mean * mean, 0.0);
    // full-data proxy: LL + G.delta + delta.H delta / 2
    double gd = 0;
#pragma unroll
    for (int q = 0; q < kEcsMaxD / kEcsThreads; ++q) {
      const int k = threadIdx.x + q * kEcsThreads;
      if (k < d) gd += dlt[k] * (a.G[k] + 0.5 * hk[q]);
    }
    const double full = a.LL + block_sum(gd, red);
    est = full + (double)a.n * mean - 0.5 * (double)a.n * nm * var;
    w1 = nm * (1.0 + nm * mean);
    w2 = nm * nm;
  } else {
    est = (double)a.n * mean;
    w1 = 1.0;
    w2 = 0.0;
  }
  // prior: sum log N(beta_k | 0, s)
  const double s2v = a.prior_scale * a.prior_scale;
  const double cst = log(a.prior_scale) + 0.91893853320467274178;  // log s + log(2 pi) / 2
  double pr = 0;
  for (int k = threadIdx.x; k < d; k += kEcsThreads)
    pr += 0.5 * (double)beta[k] * beta[k] / s2v + cst;
  pr = block_sum(pr, red);  // also orders the part[] writes before the reads below
  if (threadIdx.x == 0) pe[c] = pr - est;
#pragma unroll
  for (int q = 0; q < kEcsMaxD / kEcsThreads; ++q) {
    const int k = threadIdx.x + q * kEcsThreads;
    if (k < d) {
      double t1 = 0, t2 = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        t1 += part[0][w][k];
        t2 += part[1][w][k] + shift[w] * part[0][w][k];
      }
      double acc = w1 * t1 - w2 * t2;
      if (a.proxy) acc += a.G[k] + hk[q];
      g[(int64_t)c * ldg + k] = (float)((double)beta[k] / s2v - acc);
    }
  }
}

// ---- the Gibbs update ---------------------------------------------------------

// jax.random.randint(key, (n,), lo, hi)[i] for int32 (two 32-bit draws from
// split(key), combined as (hi % span) * (2^32 mod span) + lo % span in
// wrapping uint32, mod span).  n = 0 means the scalar shape.
__device__ __forceinline__ int randint_at(Key key, uint32_t i, uint32_t n, bool part, int lo,
                                          int hi) {
  const Key k1 = split_key(key, 0, 2, part), k2 = split_key(key, 1, 2, part);
  const uint32_t bh = n ? bits_vec(k1, i, n, part) : bits_scalar(k1, part);
  const uint32_t bl = n ? bits_vec(k2, i, n, part) : bits_scalar(k2, part);
  const uint32_t span = hi > lo ? (uint32_t)(hi - lo) : 1u;
  uint32_t mult = 65536u % span;
  mult = (mult * mult) % span;
  const uint32_t off = ((bh % span) * mult + bl % span) % span;
  return lo + (int)off;
}

// Per chain: advance the state key (split(key)[0]), pick a block with
// split(key', 3)[2] and redraw its indices with split(key', 3)[1].
__global__ __launch_bounds__(kEcsThreads) void ecs_propose_kernel(int C, int m, int nb, int64_t n,
                                                                  bool part, uint32_t* keys,
                                                                  const int* idx, int* prop) {
  const int c = blockIdx.x;
  if (c >= C) return;
  const Key k0{keys[2 * c], keys[2 * c + 1]};
  const Key rk = split_key(k0, 0, 2, part);
  const Key sub = split_key(rk, 1, 3, part), blk = split_key(rk, 2, 3, part);
  const int bs = (m - 1) / nb + 1;
  const int pad = bs - (m - 1) % bs - 1;
  int start = randint_at(blk, 0, 0, part, 0, nb) * bs;
  start = min(start, m + pad - bs);
  const int* ic = idx + (int64_t)c * m;
  int* pc = prop + (int64_t)c * m;
  for (int j = threadIdx.x; j < m; j += kEcsThreads) {
    const int t = j - start;
    pc[j] = (t >= 0 && t < bs) ? randint_at(sub, t, bs, part, 0, (int)n) : ic[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) { keys[2 * c] = rk.k0; keys[2 * c + 1] = rk.k1; }
}

// Per chain: accept the proposal with probability min(1, exp(pe - pe_new))
// by bernoulli(key', p); on acceptance the indices, the potential energy and
// the gradient (gst, the sampler's state gradient) take the proposal's values.
__global__ __launch_bounds__(kEcsThreads) void ecs_accept_kernel(
    int C, int m, int D, bool part, const uint32_t* keys, const double* pe_new, const float* g_new,
    int ldg, const int* prop, int* idx, ChainState* st, float* gst, int ldv, float* accept_out) {
  const int c = blockIdx.x;
  if (c >= C) return;
  const Key rk{keys[2 * c], keys[2 * c + 1]};
  const double pe = st[c].pe;
  // clip(exp(pe - pe_new), None, 1) keeps a NaN (a NaN proposal energy),
  // and bernoulli(key, NaN) is false: such a proposal is rejected
  const double x = exp(pe - pe_new[c]);
  const double p = x > 1.0 ? 1.0 : x;
  const bool acc = bernoulli(rk, p, part);
  if (acc) {
    for (int j = threadIdx.x; j < m; j += kEcsThreads)
      idx[(int64_t)c * m + j] = prop[(int64_t)c * m + j];
    float* gv = gst + (size_t)c * ldv;
    for (int k = threadIdx.x; k < D; k += kEcsThreads) gv[k] = g_new[(int64_t)c * ldg + k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (acc) st[c].pe = pe_new[c];
    if (accept_out) accept_out[c] = (float)p;
  }
}

__global__ void randint_kernel(const uint32_t* keys, int nk, int n, int lo, int hi, bool part,
                               int* out) {
  const int c = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nk || i >= n) return;
  out[(int64_t)c * n + i] = randint_at(Key{keys[2 * c], keys[2 * c + 1]}, i, n, part, lo, hi);
}

__global__ void split_many_kernel(const uint32_t* keys, int nk, int num, bool part,
                                  uint32_t* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nk * num) return;
  const int c = t / num, j = t % num;
  const Key k = split_key(Key{keys[2 * c], keys[2 * c + 1]}, j, num, part);
  out[2 * t] = k.k0;
  out[2 * t + 1] = k.k1;
}

struct EcsLogistic : Potential {
  EcsArgs a{};
  int C = 0, nb = 1, Dp = 0;
  bool part = true;
  int* idx = nullptr;     // caller-owned [C, m]
  uint32_t* keys = nullptr;  // caller-owned [C, 2]
  int* prop = nullptr;
  double* pe_new = nullptr;
  float* g_new = nullptr;
  double* buf = nullptr;  // bref, per-row scalars, G, H
  ~EcsLogistic() override {
    (void)hipFree(prop);
    (void)hipFree(pe_new);
    (void)hipFree(g_new);
    (void)hipFree(buf);
  }
  int dim() const override { return a.d; }
  const char* name() const override { return "ecs_eval_kernel"; }
  static_assert(kEcsMaxD == 8 * NAX_WAVE, "the widest instantiation holds 8 column chunks");
  // per chain: each subsample row once (4d), its label and index, and the
  // four fp64 proxy scalars
  double alg_bytes(int C_) const override {
    return (double)C_ * a.m * ((double)a.d * 4 + 4 + 4 + (a.proxy ? 32 : 0));
  }
  double alg_flops(int C_) const override { return 4.0 * C_ * a.m * a.d; }
  hipError_t launch(hipStream_t s, const int* ix, const float* z, int ldz, int grid, const int* act,
                    const int* nact, double* pe, float* g, int ldg, bool mark = true) {
    if (mark) mark_begin(s);
    const dim3 gr(grid), bl(kEcsThreads);
    if (a.d <= 64)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(ecs_eval_kernel<1, 8>), gr, bl, 0, s, a, ix, z, ldz, act, nact, pe, g, ldg);
    else if (a.d <= 128)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(ecs_eval_kernel<2, 8>), gr, bl, 0, s, a, ix, z, ldz, act, nact, pe, g, ldg);
    else if (a.d <= 256)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(ecs_eval_kernel<4, 4>), gr, bl, 0, s, a, ix, z, ldz, act, nact, pe, g, ldg);
    else
      hipLaunchKernelGGL(HIP_KERNEL_NAME(ecs_eval_kernel<8, 2>), gr, bl, 0, s, a, ix, z, ldz, act, nact, pe, g, ldg);
    if (mark) mark_end(s);
    return hipGetLastError();
  }
  hipError_t eval(hipStream_t s, const float* z, int ldz, int C_, double* pe, float* g,
                  int ldg) override {
    return launch(s, idx, z, ldz, C_, nullptr, nullptr, pe, g, ldg);
  }
  hipError_t eval_active(hipStream_t s, const float* z, int ldz, int C_, const int* act,
                         const int* nact, double* pe, float* g, int ldg) override {
    return launch(s, idx, z, ldz, C_, act, nact, pe, g, ldg);
  }
  hipError_t ecs_gibbs(hipStream_t s, ChainState* st, const float* z, float* gst, int ldv,
                       float* accept) override {
    hipLaunchKernelGGL(ecs_propose_kernel, dim3(C), dim3(kEcsThreads), 0, s, C, a.m, nb, a.n,
                       part, keys, idx, prop);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    // U and dU at each chain's current position with the proposed indices
    err = launch(s, prop, z, ldv, C, nullptr, nullptr, pe_new, g_new, Dp, false);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(ecs_accept_kernel, dim3(C), dim3(kEcsThreads), 0, s, C, a.m, a.d, part,
                       keys, pe_new, g_new, Dp, prop, idx, st, gst, ldv, accept);
    return hipGetLastError();
  }
};

Potential* make_ecs_logistic(const float* x, const float* y, int64_t n, int d, double prior_scale,
                             int m, int num_blocks, int proxy, const double* ref, int* idx,
                             uint32_t* keys, int num_chains, int Dp, bool partitionable,
                             hipStream_t s, std::string* err) {
  if (n <= 0 || d <= 0 || d > kEcsMaxD) { *err = "ecs model: need n > 0 and 0 < dim <= 512"; return nullptr; }
  if (m <= 0 || m > kEcsMaxM || m > n) {
    *err = "ecs model: subsample_size must be in [1, min(n, 4096)]";
    return nullptr;
  }
  if (num_blocks <= 0) { *err = "ecs model: num_blocks must be positive"; return nullptr; }
  if (proxy < 0 || proxy > 2) { *err = "ecs model: proxy degree must be 0 (none), 1 or 2"; return nullptr; }
  if (proxy && !ref) { *err = "ecs model: a Taylor proxy needs reference coefficients"; return nullptr; }
  if (!(prior_scale > 0)) { *err = "prior_scale must be positive"; return nullptr; }
  if (!idx || !keys) { *err = "ecs model: null index or key buffer"; return nullptr; }
  if (n > (int64_t)INT32_MAX) { *err = "ecs model: n must fit int32 indices"; return nullptr; }
  auto* p = new EcsLogistic();
  EcsArgs& a = p->a;
  a.x = x; a.y = y; a.n = n; a.d = d; a.m = m; a.proxy = proxy; a.prior_scale = prior_scale;
  p->C = num_chains; p->nb = num_blocks; p->Dp = Dp; p->part = partitionable;
  p->idx = idx; p->keys = keys;
  auto fail = [&](const char* what, hipError_t e) -> Potential* {
    *err = std::string(what) + ": " + hipGetErrorString(e);
    delete p;
    return nullptr;
  };
  hipError_t e;
  if ((e = hipMalloc(&p->prop, sizeof(int) * (size_t)num_chains * m))) return fail("hipMalloc", e);
  if ((e = hipMalloc(&p->pe_new, sizeof(double) * num_chains))) return fail("hipMalloc", e);
  if ((e = hipMalloc(&p->g_new, sizeof(float) * (size_t)num_chains * Dp))) return fail("hipMalloc", e);
  if (!proxy) return p;
  // [rows: n records of 4 doubles][bref d][G d][H d*d]
  const size_t nbuf = 4 * (size_t)n + d + d + (size_t)d * d;
  if ((e = hipMalloc(&p->buf, sizeof(double) * nbuf))) return fail("hipMalloc", e);
  EcsRow* rows = reinterpret_cast<EcsRow*>(p->buf);
  double* bref = p->buf + 4 * (size_t)n;
  a.rows = rows;
  a.bref = bref;
  double* G = bref + d;
  double* H = G + d;
  a.G = G; a.H = H;
  if ((e = hipMemcpyAsync(bref, ref, sizeof(double) * d, hipMemcpyDeviceToDevice, s)))
    return fail("hipMemcpyAsync", e);
  hipLaunchKernelGGL(ecs_rows_kernel, dim3(2048), dim3(kEcsThreads), 0, s, a, rows);
  if ((e = hipGetLastError())) return fail("ecs_rows_kernel", e);
  double* part = nullptr;
  const size_t npart = (size_t)kEcsSplit * ((size_t)d * d > (size_t)d + 1 ? (size_t)d * d : d + 1);
  if ((e = hipMallocAsync((void**)&part, sizeof(double) * npart, s))) return fail("hipMallocAsync", e);
  hipLaunchKernelGGL(ecs_grad_sum_kernel, dim3((d + 1 + kEcsThreads - 1) / kEcsThreads, kEcsSplit),
                     dim3(kEcsThreads), 0, s, a, part);
  double* gl = nullptr;
  if ((e = hipMallocAsync((void**)&gl, sizeof(double) * (d + 1), s))) return fail("hipMallocAsync", e);
  hipLaunchKernelGGL(ecs_reduce_kernel, dim3((d + 1 + 255) / 256), dim3(256), 0, s, part,
                     kEcsSplit, (int64_t)d + 1, gl);
  double LL = 0;
  if ((e = hipMemcpyAsync(G, gl, sizeof(double) * d, hipMemcpyDeviceToDevice, s))) return fail("copy", e);
  if ((e = hipMemcpyAsync(&LL, gl + d, sizeof(double), hipMemcpyDeviceToHost, s))) return fail("copy", e);
  if (proxy == 2) {
    hipLaunchKernelGGL(ecs_hess_kernel, dim3((d + 63) / 64, (d + 63) / 64, kEcsSplit),
                       dim3(kEcsThreads), 0, s, a, part);
    hipLaunchKernelGGL(ecs_reduce_kernel, dim3((unsigned)(((int64_t)d * d + 255) / 256)), dim3(256),
                       0, s, part, kEcsSplit, (int64_t)d * d, H);
  } else {
    if ((e = hipMemsetAsync(H, 0, sizeof(double) * d * d, s))) return fail("hipMemsetAsync", e);
  }
  if ((e = hipGetLastError())) return fail("ecs set-up", e);
  const hipError_t f1 = hipFreeAsync(part, s), f2 = hipFreeAsync(gl, s);
  if ((e = f1 != hipSuccess ? f1 : f2)) return fail("hipFreeAsync", e);
  if ((e = hipStreamSynchronize(s))) return fail("ecs set-up", e);
  a.LL = LL;
  return p;
}

}  // namespace nax

// random.randint(keys[c], (n,), lo, hi) for each of nk keys: out [nk, n].
extern "C" int nax_randint(const uint32_t* keys, int32_t nk, int32_t n, int32_t lo, int32_t hi,
                           int32_t partitionable, int32_t* out) {
  if (!keys || !out || nk <= 0 || n <= 0) {
    nax::set_error("nax_randint: null buffer or empty shape");
    return NAX_EINVAL;
  }
  hipLaunchKernelGGL(nax::randint_kernel, dim3((n + 255) / 256, nk), dim3(256), 0, nullptr, keys,
                     nk, n, lo, hi, partitionable != 0, out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    nax::set_error(std::string("nax_randint: ") + hipGetErrorString(e));
    return NAX_EHIP;
  }
  return NAX_OK;
}

// random.split(keys[c], num) for each of nk keys: out [nk, num, 2].
extern "C" int nax_split_keys(const uint32_t* keys, int32_t nk, int32_t num,
                              int32_t partitionable, uint32_t* out) {
  if (!keys || !out || nk <= 0 || num <= 0) {
    nax::set_error("nax_split_keys: null buffer or empty shape");
    return NAX_EINVAL;
  }
  const int64_t tot = (int64_t)nk * num;
  hipLaunchKernelGGL(nax::split_many_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     nullptr, keys, nk, num, partitionable != 0, out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    nax::set_error(std::string("nax_split_keys: ") + hipGetErrorString(e));
    return NAX_EHIP;
  }
  return NAX_OK;
}
