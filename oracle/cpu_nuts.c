/* CPU restatement of the logistic-regression NUTS path -- TEST AND CPU-BASELINE
 * INFRASTRUCTURE ONLY (the oracle's C leg, timed as the CPU baseline).
 *
 * Same algorithm as the float64 numpy oracle (oracle/hmc_util.py,
 * oracle/hmc.py), which restates numpyro/infer/hmc_util.py and
 * numpyro/infer/hmc.py of numpyro 0.21.0:
 *   - iterative NUTS tree with checkpointed U-turn checks
 *     (hmc_util.py:907-1180), uniform / biased multinomial transitions
 *     (hmc_util.py:749-848), divergence at dE > 1000 (hmc.py:188);
 *   - dual averaging + diagonal Welford on the Stan window schedule
 *     (hmc_util.py:60-239, 387-707);
 *   - jax.random threefry2x32 key splits / normals / bernoullis in the
 *     partitionable derivation;
 *   - potential: logistic likelihood + Normal(0, s) prior with the analytic
 *     gradient X^T(sigmoid(Xb) - y) + b/s^2 in one fused pass over X.
 * Data fp32, vectors fp64, energies fp64.  Chains run in parallel with
 * OpenMP (one chain per thread at a time).
 */
#include <alloca.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { uint32_t k0, k1; } key2;

static inline uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

static void tf(key2 k, uint32_t* a, uint32_t* b) {
  static const int R[2][4] = {{13, 15, 26, 6}, {17, 29, 16, 24}};
  uint32_t ks[3] = {k.k0, k.k1, k.k0 ^ k.k1 ^ 0x1BD11BDAu};
  uint32_t x0 = *a + ks[0], x1 = *b + ks[1];
  for (int g = 0; g < 5; ++g) {
    for (int i = 0; i < 4; ++i) { x0 += x1; x1 = rotl(x1, R[g & 1][i]); x1 ^= x0; }
    x0 += ks[(g + 1) % 3];
    x1 += ks[(g + 2) % 3] + (uint32_t)(g + 1);
  }
  *a = x0; *b = x1;
}

static key2 split_key(key2 k, uint32_t idx) {
  uint32_t a = 0, b = idx;
  tf(k, &a, &b);
  key2 r = {a, b};
  return r;
}

static float unit_f32(uint32_t bits) {
  uint32_t u = (bits >> 9) | 0x3F800000u;
  float f;
  memcpy(&f, &u, 4);
  return f - 1.0f;
}

static int bernoulli(key2 k, double p) {
  uint32_t a = 0, b = 0;
  tf(k, &a, &b);
  return (double)unit_f32(a ^ b) < p;
}

/* Giles' single-precision erfinv (the same published method the GPU uses). */
static float erfinv_f(float x) {
  float w = -log1pf(-x * x), p;
  if (w < 5.0f) {
    w = w - 2.5f;
    p = 2.81022636e-08f; p = fmaf(p, w, 3.43273939e-07f); p = fmaf(p, w, -3.5233877e-06f);
    p = fmaf(p, w, -4.39150654e-06f); p = fmaf(p, w, 0.00021858087f);
    p = fmaf(p, w, -0.00125372503f); p = fmaf(p, w, -0.00417768164f);
    p = fmaf(p, w, 0.246640727f); p = fmaf(p, w, 1.50140941f);
  } else {
    w = sqrtf(w) - 3.0f;
    p = -0.000200214257f; p = fmaf(p, w, 0.000100950558f); p = fmaf(p, w, 0.00134934322f);
    p = fmaf(p, w, -0.00367342844f); p = fmaf(p, w, 0.00573950773f);
    p = fmaf(p, w, -0.0076224613f); p = fmaf(p, w, 0.00943887047f);
    p = fmaf(p, w, 1.00167406f); p = fmaf(p, w, 2.83297682f);
  }
  return p * x;
}

static float normal_i(key2 k, uint32_t i) {
  uint32_t a = 0, b = i;
  tf(k, &a, &b);
  float f = unit_f32(a ^ b);
  const float lo = -0.99999994f;
  float u = fmaxf(lo, f * 2.0f + lo);
  return 1.41421356237309515f * erfinv_f(u);
}

typedef struct {
  const float* X; const float* y; int64_t N; int D; double inv_s2, log_norm;
} model_t;

/* U and dU in one pass over X (fp32 data, fp64 accumulation). */
static double potential_grad(const model_t* m, const double* z, double* g) {
  const int D = m->D;
  double u = 0.0;
  for (int d = 0; d < D; ++d) g[d] = z[d] * m->inv_s2;
  float* zf = (float*)alloca(sizeof(float) * D);
  float* gf = (float*)alloca(sizeof(float) * D);
  for (int d = 0; d < D; ++d) { zf[d] = (float)z[d]; gf[d] = 0.0f; }
  for (int64_t n = 0; n < m->N; ++n) {
    const float* x = m->X + n * D;
    float l = 0.0f;
    for (int d = 0; d < D; ++d) l += x[d] * zf[d];
    float yv = m->y[n];
    float r = 1.0f / (1.0f + expf(-l)) - yv;
    u += (double)(fmaxf(l, 0.0f) + log1pf(expf(-fabsf(l))) - yv * l);
    for (int d = 0; d < D; ++d) gf[d] += r * x[d];
  }
  for (int d = 0; d < D; ++d) {
    g[d] += gf[d];
    u += 0.5 * z[d] * z[d] * m->inv_s2 + m->log_norm;
  }
  return u;
}

typedef struct {
  int D, maxd;
  double *z, *g, *zl, *rl, *gl, *zr, *rr, *gr, *zp, *gp, *rs, *szp, *sgp, *srs, *imm, *msq;
  double *ckr, *cks, *wmean, *wm2;
} chain_t;

static double kinetic(const chain_t* c, const double* r) {
  double a = 0.0;
  for (int d = 0; d < c->D; ++d) a += c->imm[d] * r[d] * r[d];
  return 0.5 * a;
}

static int is_turning(const chain_t* c, const double* rl, const double* rr, const double* rs,
                      const double* rs_ck, const double* r_ck) {
  double a = 0.0, b = 0.0;
  for (int d = 0; d < c->D; ++d) {
    double s = rs[d];
    if (rs_ck) s = s - rs_ck[d] + r_ck[d];
    double q = s - 0.5 * (rl[d] + rr[d]);
    a += c->imm[d] * rl[d] * q;
    b += c->imm[d] * rr[d] * q;
  }
  return (a <= 0.0) || (b <= 0.0);
}

static double lae(double a, double b) {
  double m = a > b ? a : b;
  if (m == -INFINITY) return -INFINITY;
  return m + log1p(exp(-fabs(a - b)));
}

/* Runs C chains; returns 0.  z_out [C, num_samples, D]; num_steps_out
 * [C, num_warmup + num_samples]; *grad_evals = total leapfrog steps. */
int cpu_nuts_logistic(const float* X, const float* y, int64_t N, int D, double prior_scale,
                      int C, const float* z0, const uint32_t* keys, int num_warmup,
                      int num_samples, int max_depth, double step_size0, int adapt,
                      float* z_out, int32_t* num_steps_out, int64_t* grad_evals, int threads) {
  model_t m = {X, y, N, D, 1.0 / (prior_scale * prior_scale),
               log(sqrt(2.0 * M_PI) * prior_scale)};
  /* Stan window schedule (hmc_util.py:387-436) */
  int ends[64], nw = 0;
  if (num_warmup > 0 && num_warmup < 20) ends[nw++] = num_warmup - 1;
  else if (num_warmup >= 20) {
    int sb = 75, eb = 50, iw = 25;
    if (sb + eb + iw > num_warmup) { sb = (int)(0.15 * num_warmup); eb = (int)(0.1 * num_warmup); iw = num_warmup - sb - eb; }
    ends[nw++] = sb - 1;
    int es = num_warmup - eb, ns = iw, st = sb;
    while (st < es && nw < 62) {
      int cs = st, csz = ns;
      if (3 * csz <= es - cs) ns = 2 * csz; else csz = es - cs;
      st = cs + csz;
      ends[nw++] = st - 1;
    }
    ends[nw++] = num_warmup - 1;
  }
  int64_t total_evals = 0;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total_evals)
#endif
  for (int ci = 0; ci < C; ++ci) {
    chain_t c;
    c.D = D;
    c.maxd = max_depth;
    double* buf = (double*)calloc((size_t)D * (18 + 2 * max_depth), sizeof(double));
    double** slots[] = {&c.z, &c.g, &c.zl, &c.rl, &c.gl, &c.zr, &c.rr, &c.gr, &c.zp, &c.gp,
                        &c.rs, &c.szp, &c.sgp, &c.srs, &c.imm, &c.msq, &c.wmean, &c.wm2};
    for (int s = 0; s < 18; ++s) *slots[s] = buf + (size_t)s * D;
    c.ckr = buf + (size_t)18 * D;
    c.cks = c.ckr + (size_t)max_depth * D;
    double* rtmp = (double*)calloc(D, sizeof(double));
    for (int d = 0; d < D; ++d) { c.z[d] = z0[(size_t)ci * D + d]; c.imm[d] = 1.0; c.msq[d] = 1.0; }
    key2 k = {keys[2 * ci], keys[2 * ci + 1]};
    key2 rng = split_key(k, 0), kwa = split_key(k, 1);
    key2 akey = split_key(kwa, 0);
    double pe = potential_grad(&m, c.z, c.g);
    double step = step_size0, da_x = 0, da_xavg = 0, da_gavg = 0, da_prox = log(10.0 * step_size0);
    int da_t = 0, widx = 0, wn = 0;
    int64_t evals = 0;
    for (int it = 0; it < num_warmup + num_samples; ++it) {
      key2 kn = split_key(rng, 0), kmom = split_key(rng, 1), ktr = split_key(rng, 2);
      rng = kn;
      for (int d = 0; d < D; ++d) c.rl[d] = c.msq[d] * normal_i(kmom, d);
      memcpy(c.zl, c.z, D * 8); memcpy(c.gl, c.g, D * 8);
      memcpy(c.zr, c.z, D * 8); memcpy(c.rr, c.rl, D * 8); memcpy(c.gr, c.g, D * 8);
      memcpy(c.zp, c.z, D * 8); memcpy(c.gp, c.g, D * 8); memcpy(c.rs, c.rl, D * 8);
      double e0 = pe + kinetic(&c, c.rl), peP = pe, weight = 0, sum_acc = 0;
      int depth = 0, turning = 0, div = 0, nprop = 0;
      key2 tk = ktr;
      while (depth < max_depth && !turning && !div) {
        key2 tkn = split_key(tk, 0), kdir = split_key(tk, 1), kdbl = split_key(tk, 2);
        tk = tkn;
        int right = bernoulli(kdir, 0.5);
        key2 sk = split_key(kdbl, 0), ktop = split_key(kdbl, 1);
        double* ez = right ? c.zr : c.zl;
        double* er = right ? c.rr : c.rl;
        double* eg = right ? c.gr : c.gl;
        double eps = right ? step : -step;
        int snum = 0, sturn = 0, sdiv = 0;
        double sw = 0, ssum = 0, speP = 0;
        while (snum < (1 << depth) && !sturn && !sdiv) {
          key2 skn = split_key(sk, 0), lk = split_key(sk, 1);
          sk = skn;
          for (int d = 0; d < D; ++d) {
            er[d] -= 0.5 * eps * eg[d];
            ez[d] += eps * c.imm[d] * er[d];
          }
          double pen = potential_grad(&m, ez, eg);
          for (int d = 0; d < D; ++d) er[d] -= 0.5 * eps * eg[d];
          evals++;
          double en = pen + kinetic(&c, er);
          double de = en - e0;
          if (isnan(de)) de = INFINITY;
          double lw = -de;
          int ldiv = de > 1000.0;
          double acc = de <= 0 ? 1.0 : exp(-de);
          int li = snum;
          if (li == 0) {
            memcpy(c.szp, ez, D * 8); memcpy(c.sgp, eg, D * 8); memcpy(c.srs, er, D * 8);
            speP = pen; sw = lw; sdiv = ldiv; ssum = acc; snum = 1;
          } else {
            for (int d = 0; d < D; ++d) c.srs[d] += er[d];
            double dw = lw - sw;
            double p = dw >= 0 ? 1.0 / (1.0 + exp(-dw)) : exp(dw) / (1.0 + exp(dw));
            if (bernoulli(lk, p)) { memcpy(c.szp, ez, D * 8); memcpy(c.sgp, eg, D * 8); speP = pen; }
            sw = lae(sw, lw); sdiv = ldiv; ssum += acc; snum += 1;
          }
          int imax = __builtin_popcount((unsigned)li >> 1);
          int nsub = __builtin_popcount((unsigned)((~li & (li + 1)) - 1));
          int imin = imax - nsub + 1;
          if ((li & 1) == 0) {
            memcpy(c.ckr + (size_t)imax * D, er, D * 8);
            memcpy(c.cks + (size_t)imax * D, c.srs, D * 8);
          }
          int tn = 0;
          for (int i = imax; i >= imin && !tn; --i)
            tn = is_turning(&c, c.ckr + (size_t)i * D, er, c.srs, c.cks + (size_t)i * D, c.ckr + (size_t)i * D);
          sturn = tn;
        }
        for (int d = 0; d < D; ++d) c.rs[d] += c.srs[d];
        double p = 0.0;
        if (!sturn && !sdiv) { double t = sw - weight; p = t > 0 ? 1.0 : exp(t); }
        int tfull = sturn || is_turning(&c, c.rl, c.rr, c.rs, NULL, NULL);
        if (bernoulli(ktop, p)) { memcpy(c.zp, c.szp, D * 8); memcpy(c.gp, c.sgp, D * 8); peP = speP; }
        depth++; weight = lae(weight, sw); turning = tfull; div = sdiv; sum_acc += ssum; nprop += snum;
      }
      double ap = sum_acc / nprop;
      memcpy(c.z, c.zp, D * 8); memcpy(c.g, c.gp, D * 8);
      pe = peP;
      if (num_steps_out) num_steps_out[(size_t)ci * (num_warmup + num_samples) + it] = nprop;
      if (it < num_warmup && adapt) {
        akey = split_key(akey, 0);
        da_t++;
        double gv = 0.8 - ap;
        da_gavg = (1.0 - 1.0 / (da_t + 10.0)) * da_gavg + gv / (da_t + 10.0);
        da_x = da_prox - sqrt((double)da_t) / 0.05 * da_gavg;
        double w = pow((double)da_t, -0.75);
        da_xavg = (1.0 - w) * da_xavg + w * da_x;
        step = (it == num_warmup - 1) ? exp(da_xavg) : exp(da_x);
        int mid = (0 < widx) && (widx < nw - 1);
        if (mid) {
          wn++;
          for (int d = 0; d < D; ++d) {
            double dp = c.z[d] - c.wmean[d];
            c.wmean[d] += dp / wn;
            c.wm2[d] += dp * (c.z[d] - c.wmean[d]);
          }
        }
        int at_end = it == ends[widx];
        if (at_end) widx++;
        if (at_end && mid) {
          for (int d = 0; d < D; ++d) {
            double cov = c.wm2[d] / (wn - 1.0);
            cov = (wn / (wn + 5.0)) * cov + 1e-3 * (5.0 / (wn + 5.0));
            c.imm[d] = cov; c.msq[d] = 1.0 / sqrt(cov); c.wmean[d] = 0; c.wm2[d] = 0;
          }
          wn = 0;
          da_x = 0; da_xavg = 0; da_gavg = 0; da_t = 0; da_prox = log(10.0) + log(step);
        }
      }
      if (it >= num_warmup && z_out)
        for (int d = 0; d < D; ++d)
          z_out[((size_t)ci * num_samples + (it - num_warmup)) * D + d] = (float)c.z[d];
    }
    total_evals += evals;
    free(rtmp);
    free(buf);
  }
  if (grad_evals) *grad_evals = total_evals;
  return 0;
}

/* The vectorized chain method's per-leapfrog work on the CPU (mcmc.py:720-724
 * vmaps the chains, so each leapfrog step of the batch is one X Z^T and one
 * X^T R over all C chains): U and dU/dz of the logistic model for C chains at
 * z [C][D], with every thread taking a contiguous block of rows and passing
 * over each of its rows once for all chains (logits for all chains, then the
 * rank-1 update of its private gradient block).  Thread partials are summed
 * in thread order.  U [C] fp64, g [C][D] fp64.  Returns 0. */
/* fast-math for this function only: lets gcc call the vector expf / log1pf
 * of glibc's libmvec for the per-(row, chain) terms */
__attribute__((optimize("fast-math")))
int cpu_logistic_batch_grad(const float* X, const float* y, int64_t N, int D, double prior_scale,
                            int Cn, const float* z, double* U, double* g, int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
  const int T = threads > 0 ? threads : omp_get_max_threads();
#else
  const int T = 1;
#endif
  const int Cp = (Cn + 15) / 16 * 16;                              /* chains padded to 16 */
  float* zt = (float*)calloc((size_t)D * Cp, sizeof(float));       /* [D][Cp] */
  float* gp = (float*)calloc((size_t)T * D * Cp, sizeof(float));  /* per-thread [D][Cp] */
  double* up = (double*)calloc((size_t)T * Cn, sizeof(double));
  for (int c = 0; c < Cn; ++c)
    for (int d = 0; d < D; ++d) zt[(size_t)d * Cp + c] = z[(size_t)c * D + d];
#pragma omp parallel num_threads(T)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    const int64_t r0 = N * t / T, r1 = N * (t + 1) / T;
    float* gt = gp + (size_t)t * D * Cp;
    double* ut = up + (size_t)t * Cn;
    /* four rows at a time, chains in blocks of 16 held in registers across
     * the d loop: each zt / gradient row is loaded once per 4 rows */
    float* restrict l = (float*)malloc(sizeof(float) * 4 * Cp);
    for (int64_t n0 = r0; n0 < r1; n0 += 4) {
      const int nr = (int)(r1 - n0 < 4 ? r1 - n0 : 4);
      const float* xq[4];
      for (int q = 0; q < 4; ++q) xq[q] = X + (n0 + (q < nr ? q : 0)) * D;
      for (int c0 = 0; c0 < Cp; c0 += 16) {
        float acc[4][16];
        for (int q = 0; q < 4; ++q)
          for (int j = 0; j < 16; ++j) acc[q][j] = 0.0f;
        for (int d = 0; d < D; ++d) {
          const float* restrict zd = zt + (size_t)d * Cp + c0;
          for (int q = 0; q < 4; ++q) {
            const float xv = xq[q][d];
            for (int j = 0; j < 16; ++j) acc[q][j] += xv * zd[j];
          }
        }
        for (int q = 0; q < 4; ++q)
          for (int j = 0; j < 16; ++j) l[(size_t)q * Cp + c0 + j] = acc[q][j];
      }
      for (int q = 0; q < nr; ++q) {
        const float yv = y[n0 + q];
        float* restrict lq = l + (size_t)q * Cp;
        for (int c = 0; c < Cn; ++c) {
          const float lc = lq[c];
          ut[c] += (double)(fmaxf(lc, 0.0f) + log1pf(expf(-fabsf(lc))) - yv * lc);
          lq[c] = 1.0f / (1.0f + expf(-lc)) - yv;
        }
        for (int c = Cn; c < Cp; ++c) lq[c] = 0.0f;
      }
      for (int q = nr; q < 4; ++q)
        for (int c = 0; c < Cp; ++c) l[(size_t)q * Cp + c] = 0.0f;
      for (int d = 0; d < D; ++d) {
        float* restrict gd = gt + (size_t)d * Cp;
        const float x0 = xq[0][d], x1 = xq[1][d], x2 = xq[2][d], x3 = xq[3][d];
        const float* restrict l0 = l;
        const float* restrict l1 = l + Cp;
        const float* restrict l2 = l + 2 * (size_t)Cp;
        const float* restrict l3 = l + 3 * (size_t)Cp;
        for (int c = 0; c < Cp; ++c) gd[c] += x0 * l0[c] + x1 * l1[c] + x2 * l2[c] + x3 * l3[c];
      }
    }
    free(l);
  }
  const double inv_s2 = 1.0 / (prior_scale * prior_scale);
  const double lnorm = log(sqrt(2.0 * M_PI) * prior_scale);
  for (int c = 0; c < Cn; ++c) {
    double u = 0.0;
    for (int t = 0; t < T; ++t) u += up[(size_t)t * Cn + c];
    for (int d = 0; d < D; ++d) {
      const double zv = z[(size_t)c * D + d];
      double s = 0.0;
      for (int t = 0; t < T; ++t) s += gp[((size_t)t * D + d) * Cp + c];
      g[(size_t)c * D + d] = s + zv * inv_s2;
      u += 0.5 * zv * zv * inv_s2 + lnorm;
    }
    U[c] = u;
  }
  free(zt);
  free(gp);
  free(up);
  return 0;
}
