// The NUTS / HMC per-chain state machine as inline device functions, shared
// by the tick kernel (sampler.hip) and the persistent per-model kernels.
//
// One wavefront owns one chain, or (persistent launches of latency-bound
// models) a group of WPC wavefronts of one workgroup does: vector passes
// then stride over the group's 64 WPC threads, reductions and scans combine
// the waves through LDS, and every wave runs the scalar control flow on its
// own copy of the chain state (the copies stay identical: every reduction
// hands all threads the same value).  advance_chain() runs the chain's control flow
// -- sample_kernel (numpyro/infer/hmc.py:459-530), build_tree
// (numpyro/infer/hmc_util.py:1088-1180) with the iterative subtree, the
// checkpointed U-turn checks and the uniform/biased transitions
// (hmc_util.py:710-1085), fixed-length HMC (hmc.py:364-414) and the warmup
// adapter (hmc_util.py:518-707) -- until the chain needs its next gradient
// (returns true, the position is in zr) or has done `target` iterations
// (returns false).  gi / pe_in carry the gradient and energy requested by the
// previous call.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "nax_device.h"
#include "nax_internal.h"

namespace nax {

// ---- per-chain vector slots (each [Dp] fp32) --------------------------------
// The slots every leaf touches come first (V_HOT of them): a persistent
// launch with one wavefront per workgroup keeps the leading slots in LDS.
enum {
  V_ZL = 0, V_RL, V_GL, V_ZR, V_RR, V_GR,  // tree edges (the subtree's outer edge lives here too)
  V_ZP, V_GP,                      // tree proposal
  V_RSUM,                          // tree r_sum
  V_SZP, V_SGP, V_SRSUM,           // subtree proposal and r_sum
  V_IMM,                           // inverse mass matrix diag
  V_HOT,
  V_Z = V_HOT, V_G,                // HMCState.z, z_grad
  V_MSQRT,                         // mass_matrix_sqrt diag
  V_R0,                            // momentum drawn by init_kernel
  V_TMP,                           // scratch vector of the dense-mass products
  V_CK                             // r_ckpts[depth], then r_sum_ckpts[depth]
};

enum {
  PH_INIT_WAIT = 0, PH_ITER_START, PH_DOUBLING, PH_LEAF, PH_LEAF_WAIT, PH_COMBINE,
  PH_ITER_END, PH_HMC_STEP, PH_HMC_WAIT, PH_HMC_END,
  PH_INIT_DONE, PH_ITER_FINISH, PH_FS_START, PH_FS_WAIT
};

__device__ __forceinline__ Key K(const uint32_t* k) { return Key{k[0], k[1]}; }
__device__ __forceinline__ void setK(uint32_t* d, Key k) { d[0] = k.k0; d[1] = k.k1; }

// A translation unit whose chains have long latent vectors (thousands of
// entries per lane group) defines NAX_FOR_D_UNROLL before including this
// header, so that several of a lane's loads are in flight at once; the
// default keeps register pressure low for short vectors.
#ifndef NAX_FOR_D_UNROLL
#define NAX_FOR_D_UNROLL 1
#endif

// LDSV: the V_HOT leading slots live in LDS (persistent launches of
// long-vector models), the others in the global vector block.  Slot numbers
// are compile-time constants or provably below / above V_HOT wherever a pass
// touches a vector, so the compiler addresses each one as LDS or as global
// memory (no generic pointers).
template <bool LDSV, int WPC_ = 1>
struct ChainCtxT {
  static constexpr int WPC = WPC_;          // wavefronts working on the chain
  static constexpr int NT = NAX_WAVE * WPC_;  // threads working on the chain
  // FOR_D's unroll: the translation unit's factor for one wave, divided by
  // the waves (each thread walks WPC times fewer elements)
  static constexpr int UNROLL = NAX_FOR_D_UNROLL / WPC_ > 0 ? NAX_FOR_D_UNROLL / WPC_ : 1;
  double* red;     // LDS reduction slots of the group (kRedSlots * WPC), WPC > 1
  float* redf;     // LDS [WPC][64] fp32 exchange of the group, WPC > 1
  float* vb;       // chain's vector block [nvec][Dp]
  float* lb;       // LDS copy of slots [0, V_HOT) when LDSV, else unused
  double* wel;     // Welford [2][Dp] (mean, m2) fp64
  int lane, D, Dp;  // lane: the thread's index in the chain's group, 0..NT-1
  // dense mass matrix ([Dp][Dp] row-major blocks, nullptr when diagonal):
  // mi[k][d] = M^-1[d][k], ms[k][d] = mass_matrix_sqrt[d][k], m2 fp64
  float* mi;
  float* ms;
  double* m2;
  __device__ float* v(int slot) const {
    if constexpr (LDSV) return (slot < V_HOT ? lb : vb) + (size_t)slot * Dp;
    return vb + (size_t)slot * Dp;
  }
};
using ChainCtx = ChainCtxT<false>;

#define NAX_PRAGMA_(x) _Pragma(#x)
#define NAX_PRAGMA(x) NAX_PRAGMA_(x)
#define NAX_CTX_T std::remove_cv_t<std::remove_reference_t<decltype(cx)>>
#define FOR_D(d) \
  NAX_PRAGMA(unroll(NAX_CTX_T::UNROLL)) for (int d = cx.lane; d < cx.D; d += cx.NT)

// Makes this wavefront's earlier global writes visible to its other lanes.
__device__ __forceinline__ void wave_sync_mem() { __threadfence_block(); }

// ---- chain-group primitives (one wave, or WPC waves of one workgroup) -------------
constexpr int kRedSlots = 8;  // values one chain_sum_n can combine

// Makes the group's earlier writes (global and LDS) visible to all its threads.
template <class G>
__device__ __forceinline__ void chain_sync(const G& g) {
  if constexpr (G::WPC == 1) __threadfence_block();
  else __syncthreads();
}

// Sums of N values over the group's threads; every thread receives the same
// sums (the waves' partials are added in wave order).
template <int N, class G>
__device__ __forceinline__ void chain_sum_n(const G& g, double (&v)[N]) {
  static_assert(N <= kRedSlots, "chain_sum_n: too many values");
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = wave_sum(v[i]);
  if constexpr (G::WPC > 1) {
    const int w = g.lane / NAX_WAVE;
    __syncthreads();  // earlier readers of the slots are done
    if ((g.lane & (NAX_WAVE - 1)) == 0)
#pragma unroll
      for (int i = 0; i < N; ++i) g.red[i * G::WPC + w] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < G::WPC; ++k) t += g.red[i * G::WPC + k];
      v[i] = t;
    }
  }
}

template <class G>
__device__ __forceinline__ double chain_sum(const G& g, double x) {
  double v[1] = {x};
  chain_sum_n(g, v);
  return v[0];
}

// Inclusive prefix sum over the group's threads in lane order.
template <class G>
__device__ __forceinline__ double chain_inclusive_scan(const G& g, double v) {
  v = wave_inclusive_scan(v, g.lane & (NAX_WAVE - 1));
  if constexpr (G::WPC > 1) {
    const int w = g.lane / NAX_WAVE;
    __syncthreads();
    if ((g.lane & (NAX_WAVE - 1)) == NAX_WAVE - 1) g.red[w] = v;
    __syncthreads();
    double off = 0.0;
    for (int k = 0; k < w; ++k) off += g.red[k];
    v += off;
  }
  return v;
}

// The group a standalone (non-persistent) evaluation of one chain runs on:
// one wavefront.
struct WaveGroup {
  static constexpr int WPC = 1;
  static constexpr int NT = NAX_WAVE;
  double* red = nullptr;
  float* redf = nullptr;
  int lane;
};

// (A x)[d] for a matrix stored transposed (at[k][d] = A[d][k]): lane d walks
// column d of `at` (coalesced across lanes) against the broadcast x[k].
template <class CX>
__device__ __forceinline__ float dense_mv(const CX& cx, const float* at, const float* x,
                                          int d) {
  float a0 = 0.0f, a1 = 0.0f;
  int k = 0;
  for (; k + 1 < cx.D; k += 2) {
    a0 = fmaf(at[(size_t)k * cx.Dp + d], x[k], a0);
    a1 = fmaf(at[(size_t)(k + 1) * cx.Dp + d], x[k + 1], a1);
  }
  if (k < cx.D) a0 = fmaf(at[(size_t)k * cx.Dp + d], x[k], a0);
  return a0 + a1;
}

template <class CX>
__device__ __forceinline__ void vcopy(const CX& cx, int dst, int src) {
  float* a = cx.v(dst);
  const float* b = cx.v(src);
  FOR_D(d) a[d] = b[d];
}

// euclidean_kinetic_energy (hmc_util.py:1183-1200): 0.5 r . (M^-1 r)
template <class CX>
__device__ __forceinline__ double kinetic(const CX& cx, int r_slot) {
  const float* r = cx.v(r_slot);
  double acc = 0.0;
  if (cx.mi) {
    chain_sync(cx);
    FOR_D(d) acc += (double)dense_mv(cx, cx.mi, r, d) * (double)r[d];
  } else {
    const float* m = cx.v(V_IMM);
    FOR_D(d) { double rv = r[d]; acc += (double)m[d] * rv * rv; }
  }
  return 0.5 * chain_sum(cx, acc);
}

// _is_turning (hmc_util.py:710-746) with r_sum given as rs - rs_ck + r_ck when
// rs_ck != nullptr (the iterative subtree form, hmc_util.py:961-981).
template <class CX>
__device__ __forceinline__ bool is_turning(const CX& cx, const float* rl, const float* rr,
                           const float* rs, const float* rs_ck, const float* r_ck) {
  const float* m = cx.v(V_IMM);
  double a = 0.0, b = 0.0;
  if (cx.mi) chain_sync(cx);
  FOR_D(d) {
    float s = rs[d];
    if (rs_ck) s = s - rs_ck[d] + r_ck[d];
    float rlv = rl[d], rrv = rr[d];
    float c = s - (rlv + rrv) * 0.5f;
    // velocities v = M^-1 r (_kinetic_grad, hmc_util.py:1203-1220)
    const float vl = cx.mi ? dense_mv(cx, cx.mi, rl, d) : m[d] * rlv;
    const float vr = cx.mi ? dense_mv(cx, cx.mi, rr, d) : m[d] * rrv;
    a += (double)vl * c;
    b += (double)vr * c;
  }
  double ab[2] = {a, b};
  chain_sum_n(cx, ab);
  return (ab[0] <= 0.0) || (ab[1] <= 0.0);
}

// momentum_generator (hmc.py:92-110), diagonal mass: s * normal(key, (D,)) with
// s = sqrt(M) (the slot V_MSQRT); find_reasonable_step_size passes the inverse
// mass matrix in that argument (hmc_util.py:359), hence the slot parameter.
// Dense mass: r = mass_matrix_sqrt . eps (hmc.py:107-109).
template <class CX>
__device__ __forceinline__ void draw_momentum(const CX& cx, int dst, Key k, bool part,
                              int scale_slot = V_MSQRT) {
  float* r = cx.v(dst);
  if (cx.mi) {
    float* eps = cx.v(V_TMP);
    FOR_D(d) eps[d] = normal_vec(k, (uint32_t)d, (uint32_t)cx.D, part);
    chain_sync(cx);
    const float* at = (scale_slot == V_IMM) ? cx.mi : cx.ms;
    FOR_D(d) r[d] = dense_mv(cx, at, eps, d);
    chain_sync(cx);
    return;
  }
  const float* s = cx.v(scale_slot);
  FOR_D(d) r[d] = s[d] * normal_vec(k, (uint32_t)d, (uint32_t)cx.D, part);
}

// First half of velocity_verlet (hmc_util.py:289-297) on an edge, in place:
// r <- r - eps/2 g; z <- z + eps M^-1 r; z is also written to the request
// (unless zreq is null: the potential then reads the edge's slot).
template <class CX>
__device__ __forceinline__ void leapfrog_half(const CX& cx, int zs, int rs, int gs, float eps,
                              float* zreq) {
  float* z = cx.v(zs);
  float* r = cx.v(rs);
  const float* g = cx.v(gs);
  const float* m = cx.v(V_IMM);
  const float h = 0.5f * eps;
  if (cx.mi) {
    FOR_D(d) r[d] = r[d] - h * g[d];
    chain_sync(cx);
    FOR_D(d) {
      const float zn = z[d] + eps * dense_mv(cx, cx.mi, r, d);
      z[d] = zn;
      if (zreq) zreq[d] = zn;
    }
    return;
  }
  FOR_D(d) {
    float rh = r[d] - h * g[d];
    float zn = z[d] + eps * (m[d] * rh);
    r[d] = rh;
    z[d] = zn;
    if (zreq) zreq[d] = zn;
  }
}

// Second half (hmc_util.py:298-309): g <- grad(z); r <- r - eps/2 g.
template <class CX>
__device__ __forceinline__ void leapfrog_finish(const CX& cx, int rs, int gs, float eps,
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

// A leaf's vector work in one pass (diagonal mass): the second half of
// velocity_verlet (g <- gin, r <- r - eps/2 g), the kinetic energy's sum
// r . M^-1 r, the subtree's r_sum (set at its first leaf, else += r), at an
// even leaf the checkpoints (r, r_sum) into ckr_w / cks_w, and at an odd leaf
// _is_turning's two sums for each of the NCHK checkpoints ck[j] (r_sum taken
// from checkpoint j on, hmc_util.py:961-981); one reduction for all of them.
// Each value is the same expression summed in the same order as by
// leapfrog_finish, kinetic and is_turning one after the other.
template <int NCHK, class CX>
__device__ __forceinline__ void leaf_pass(const CX& cx, int rs, int gs, float eps,
                                          const float* gin, bool first, float* ckr_w,
                                          float* cks_w, const float* ckr0, const float* cks0,
                                          const float* ckr1, const float* cks1,
                                          const float* ckr2, const float* cks2,
                                          double (&out)[1 + 2 * NCHK]) {
  const float* ckr[3] = {ckr0, ckr1, ckr2};
  const float* cks[3] = {cks0, cks1, cks2};
  float* r = cx.v(rs);
  float* g = cx.v(gs);
  float* srs = cx.v(V_SRSUM);
  const float* m = cx.v(V_IMM);
  const float h = 0.5f * eps;
  double acc = 0.0;
  double a[NCHK > 0 ? NCHK : 1], b[NCHK > 0 ? NCHK : 1];
#pragma unroll
  for (int j = 0; j < NCHK; ++j) a[j] = b[j] = 0.0;
  FOR_D(d) {
    const float gv = gin[d];
    g[d] = gv;
    const float rv = r[d] - h * gv;
    r[d] = rv;
    const double rd = rv;
    acc += (double)m[d] * rd * rd;
    const float s = first ? rv : srs[d] + rv;
    srs[d] = s;
    if constexpr (NCHK == 0) {
      ckr_w[d] = rv;
      cks_w[d] = s;
    }
#pragma unroll
    for (int j = 0; j < NCHK; ++j) {
      const float rlv = ckr[j][d];
      const float sj = s - cks[j][d] + rlv;
      const float c = sj - (rlv + rv) * 0.5f;
      const float vl = m[d] * rlv;
      const float vr = m[d] * rv;
      a[j] += (double)vl * c;
      b[j] += (double)vr * c;
    }
  }
  out[0] = acc;
#pragma unroll
  for (int j = 0; j < NCHK; ++j) {
    out[1 + 2 * j] = a[j];
    out[2 + 2 * j] = b[j];
  }
  chain_sum_n(cx, out);
}

// ---- warmup adaptation (hmc_util.py:60-130, 133-239, 518-707) ---------------

__device__ void da_init(ChainState& S, double prox) {
  S.da_x = 0.0; S.da_xavg = 0.0; S.da_gavg = 0.0; S.da_t = 0; S.da_prox = prox;
}

template <class CX>
__device__ void welford_reset(const CX& cx, ChainState& S) {
  double* mean = cx.wel;
  double* m2 = cx.wel + cx.Dp;
  FOR_D(d) { mean[d] = 0.0; m2[d] = 0.0; }
  S.welford_n = 0;
}

// ---- dense mass matrix (dense_mass=True) --------------------------------------
// welford_covariance update, dense (hmc_util.py:171-196): with
// dpre = x - mean, mean += dpre / n, dpost = x - mean: m2 += outer(dpost, dpre).
template <class CX>
__device__ void welford_dense_update(const CX& cx, const float* z, int n) {
  double* mean = cx.wel;
  double* dpost = cx.wel + cx.Dp;
  double* dpre = cx.wel + 2 * cx.Dp;
  FOR_D(d) {
    const double x = z[d];
    const double a = x - mean[d];
    const double mn = mean[d] + a / n;
    mean[d] = mn;
    dpre[d] = a;
    dpost[d] = x - mn;
  }
  wave_sync_mem();
  FOR_D(j) {
    const double pj = dpre[j];
    for (int i = 0; i < cx.D; ++i) cx.m2[(size_t)i * cx.Dp + j] += dpost[i] * pj;
  }
  wave_sync_mem();
}

// From a symmetric positive-definite A (fp64, in cx.m2, upper triangle read),
// the factors welford_covariance's final_fn (hmc_util.py:227-233) and
// _initialize_mass_matrix (hmc_util.py:497-507) take: the upper triangular U
// with A = U U^T (the Cholesky factor of the index-reversed matrix, reversed
// back; tril_inv = U^T) and mass_matrix_sqrt = S = (U^T)^-1, lower
// triangular.  U overwrites A's upper triangle, S's strict lower triangle its
// lower one (S's diagonal is 1 / diag U); S is then written to cx.ms
// transposed (fp32).  One wavefront, O(D^3 / 3) fp64 operations.
template <class CX>
__device__ void dense_factor(const CX& cx) {
  double* A = cx.m2;
  const int D = cx.D;
  const size_t ld = cx.Dp;
  const int lane = cx.lane;
  for (int j = D - 1; j >= 0; --j) {
    double s = 0.0;
    for (int k = j + 1 + lane; k < D; k += NAX_WAVE) {
      const double u = A[j * ld + k];
      s += u * u;
    }
    s = wave_sum(s);
    const double ujj = sqrt(A[j * ld + j] - s);
    for (int i = lane; i < j; i += NAX_WAVE) {
      double t = A[i * ld + j];
      for (int k = j + 1; k < D; ++k) t -= A[i * ld + k] * A[j * ld + k];
      A[i * ld + j] = t / ujj;
    }
    if (lane == 0) A[j * ld + j] = ujj;
    wave_sync_mem();
  }
  // S = L^-1 with L[i][k] = U[k][i]: forward substitution, lane = column.
  for (int c0 = 0; c0 < D; c0 += NAX_WAVE) {
    const int c = c0 + lane;
    const double dinv = c < D ? 1.0 / A[(size_t)c * ld + c] : 0.0;
    for (int i = c0 + 1; i < D; ++i) {
      double t = 0.0;
      for (int k = c0; k < i; ++k) {
        const double skc = (k > c) ? A[k * ld + c] : (k == c ? dinv : 0.0);
        t -= A[k * ld + i] * skc;
      }
      if (c < D && i > c) A[i * ld + c] = t / A[i * ld + i];
    }
  }
  wave_sync_mem();
  for (int k = 0; k < D; ++k)
    FOR_D(d) {
      const double v = (d > k) ? A[d * ld + k] : (d == k ? 1.0 / A[k * ld + k] : 0.0);
      cx.ms[k * ld + d] = (float)v;
    }
  wave_sync_mem();
}

template <class CX>
__device__ void dense_zero(const CX& cx) {
  for (int i = 0; i < cx.D; ++i) FOR_D(j) cx.m2[(size_t)i * cx.Dp + j] = 0.0;
  FOR_D(d) cx.wel[d] = 0.0;
  wave_sync_mem();
}

// Window end, dense: cov = m2 / (n - 1), Stan regularisation
// (hmc_util.py:215-226), M^-1 = cov, factors of cov, Welford reset.
template <class CX>
__device__ void dense_window_end(const CX& cx, double n, bool regularize) {
  const size_t ld = cx.Dp;
  for (int i = 0; i < cx.D; ++i)
    FOR_D(j) {
      double a = cx.m2[i * ld + j] / (n - 1.0);
      if (regularize) a = (n / (n + 5.0)) * a + (i == j ? 1e-3 * (5.0 / (n + 5.0)) : 0.0);
      cx.m2[i * ld + j] = a;
      cx.mi[(size_t)j * ld + i] = (float)a;
    }
  wave_sync_mem();
  dense_factor(cx);
  dense_zero(cx);
}

// warmup_adapter's update_fn (hmc_util.py:637-707).  Returns true when the
// window that ends at t is a middle window and find_reasonable_step_size must
// run now (its key is left in S.fs_key); the dual-averaging restart of
// _update_at_window_end (hmc_util.py:618-623) then happens when it returns.
template <class CX>
__device__ __forceinline__ bool adapt_update(const EngineCfg& cfg, const CX& cx, ChainState& S, int t,
                             double accept_prob) {
  const bool part = cfg.partitionable != 0;
  Key ak = K(S.adapt_key);
  setK(S.adapt_key, split_key(ak, 0, 2, part));
  setK(S.fs_key, split_key(ak, 1, 2, part));  // rng_key_ss
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
  if (cfg.adapt_mass_matrix && middle && cx.mi) {
    welford_dense_update(cx, cx.v(V_Z), S.welford_n + 1);
    S.welford_n = S.welford_n + 1;
  } else if (cfg.adapt_mass_matrix && middle) {
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
    if (cfg.adapt_mass_matrix && cx.mi) {
      dense_window_end(cx, (double)S.welford_n, cfg.regularize != 0);
      S.welford_n = 0;
    } else if (cfg.adapt_mass_matrix) {
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
    if (cfg.adapt_step_size) {
      if (cfg.find_heuristic) return true;
      da_init(S, log(10.0) + log(S.step_size));
    }
  }
  return false;
}

__device__ __forceinline__ void fs_begin(ChainState& S, int ret) {
  S.fs_step = S.step_size;
  S.fs_dir = 0;
  S.fs_last_dir = 0;
  S.fs_ret = ret;
  S.phase = PH_FS_START;
}

// ---- collection (fori_collect, numpyro/util.py:384-403) ---------------------
template <class CX>
__device__ void collect(const EngineCfg& cfg, const Collect& col, const CX& cx,
                        const ChainState& S, int c, int64_t it) {
  if (col.size <= 0) return;
  int64_t rel = it - col.start_idx;
  if (rel < 0) return;
  if ((rel % col.thinning) != col.thinning - 1) return;
  int64_t slot = rel / col.thinning;
  if (slot >= col.size) return;
  const size_t cs = (size_t)c * col.size + slot;
  if (col.z) {
    float* dst = col.z + cs * (size_t)cfg.D;
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
template <class CX>
__device__ __forceinline__ bool advance_chain(const EngineCfg& cfg, const Collect& col,
                                              int64_t target, ChainState& S,
                                              const CX& cx, int c, float* zr,
                                              const float* gi, double pe_in, int& evals) {
  const bool part = cfg.partitionable != 0;
  const int CKR = V_CK, CKS = V_CK + cfg.max_depth_alloc;
  bool need_grad = false;
  while (!need_grad) {
    switch (S.phase) {
      case PH_INIT_WAIT: {  // init_kernel (hmc.py:338-349) after value_and_grad(z0)
        float* g = cx.v(V_G);
        FOR_D(d) g[d] = gi[d];
        S.pe = pe_in;
        if (cfg.find_heuristic && cfg.adapt_step_size) fs_begin(S, PH_INIT_DONE);
        else S.phase = PH_INIT_DONE;
        break;
      }
      case PH_INIT_DONE: {  // rest of init_kernel: energy of the initial momentum
        S.energy = S.pe + kinetic(cx, V_R0);
        S.phase = PH_ITER_START;
        break;
      }
      case PH_FS_START: {  // find_reasonable_step_size's while_loop (hmc_util.py:351-383)
        const int dir = S.fs_dir, last = S.fs_last_dir;
        const bool small_ok = (S.fs_step > 2.2250738585072014e-308) || (dir >= 0);
        const bool large_ok = (S.fs_step < 1.7976931348623157e308) || (dir <= 0);
        if (!(small_ok && large_ok && (last == 0 || dir == last))) {
          S.step_size = S.fs_step;
          if (S.fs_ret == PH_INIT_DONE) da_init(S, log(10.0 * S.step_size));  // hmc_util.py:576
          else da_init(S, log(10.0) + log(S.step_size));                      // hmc_util.py:623
          S.phase = S.fs_ret;
          break;
        }
        Key fk = K(S.fs_key);
        setK(S.fs_key, split_key(fk, 0, 2, part));
        const Key kmom = split_key(fk, 1, 2, part);
        S.fs_step = (dir > 0) ? 2.0 * S.fs_step : (dir < 0 ? 0.5 * S.fs_step : S.fs_step);
        draw_momentum(cx, V_RR, kmom, part, V_IMM);
        vcopy(cx, V_ZR, V_Z);
        vcopy(cx, V_GR, V_G);
        S.fs_e0 = S.pe + kinetic(cx, V_RR);
        leapfrog_half(cx, V_ZR, V_RR, V_GR, (float)S.fs_step, zr);
        S.phase = PH_FS_WAIT;
        need_grad = true;
        break;
      }
      case PH_FS_WAIT: {
        evals++;
        leapfrog_finish(cx, V_RR, V_GR, (float)S.fs_step, gi);
        const double delta = (pe_in + kinetic(cx, V_RR)) - S.fs_e0;
        S.fs_last_dir = S.fs_dir;
        S.fs_dir = (-0.22314355131420976 < -delta) ? 1 : -1;  // log(0.8) < -delta (NaN: -1)
        S.phase = PH_FS_START;
        break;
      }
      case PH_ITER_START: {  // sample_kernel (hmc.py:459-490)
        if ((int64_t)S.i >= target) return need_grad;
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
        const int leaf_idx = S.s_num;
        // _leaf_idx_to_ckpt_idxs (hmc_util.py:941-958)
        const int idx_max = __popc((unsigned)leaf_idx >> 1);
        const int nsub = __popc((unsigned)((~leaf_idx & (leaf_idx + 1)) - 1));
        const int idx_min = idx_max - nsub + 1;
        const bool even = (leaf_idx & 1) == 0;  // even: writes checkpoint idx_max, no checks
        const float* znew = cx.v(zs);
        const float* rnew = cx.v(rs);
        const float* gnew = cx.v(gs);
        float* srs = cx.v(V_SRSUM);
        double kin;
        bool turning = false;
        int next_check = idx_max;  // U-turn checks still to make: next_check down to idx_min
        if (cx.mi) {
          leapfrog_finish(cx, rs, gs, eps, gi);
          kin = kinetic(cx, rs);
          FOR_D(d) srs[d] = leaf_idx == 0 ? rnew[d] : srs[d] + rnew[d];
          if (even) {
            float* ckr = cx.v(CKR + idx_max);
            float* cks = cx.v(CKS + idx_max);
            FOR_D(d) { ckr[d] = rnew[d]; cks[d] = srs[d]; }
          }
        } else {  // the same values from one pass and one reduction
          const int nf = even ? 0 : (nsub < 3 ? nsub : 3);
          // checkpoints idx_max, idx_max - 1, idx_max - 2 (those beyond nf unused)
          const int i1 = nf > 1 ? idx_max - 1 : idx_max, i2 = nf > 2 ? idx_max - 2 : idx_max;
          const float* r0 = cx.v(CKR + idx_max);
          const float* s0 = cx.v(CKS + idx_max);
          const float* r1 = cx.v(CKR + i1);
          const float* s1 = cx.v(CKS + i1);
          const float* r2 = cx.v(CKR + i2);
          const float* s2 = cx.v(CKS + i2);
          float* ckr_w = even ? cx.v(CKR + idx_max) : nullptr;
          float* cks_w = even ? cx.v(CKS + idx_max) : nullptr;
          const bool first = leaf_idx == 0;
          // _is_turning's verdict for checkpoint j of the pass: either sum <= 0
          auto turn = [](const double* o, int j) {
            return o[1 + 2 * j] <= 0.0 || o[2 + 2 * j] <= 0.0;
          };
          switch (nf) {
            case 0: {
              double o[1];
              leaf_pass<0>(cx, rs, gs, eps, gi, first, ckr_w, cks_w, r0, s0, r1, s1, r2, s2,
                           o);
              kin = 0.5 * o[0];
              break;
            }
            case 1: {
              double o[3];
              leaf_pass<1>(cx, rs, gs, eps, gi, first, ckr_w, cks_w, r0, s0, r1, s1, r2, s2,
                           o);
              kin = 0.5 * o[0];
              turning = turn(o, 0);
              break;
            }
            case 2: {
              double o[5];
              leaf_pass<2>(cx, rs, gs, eps, gi, first, ckr_w, cks_w, r0, s0, r1, s1, r2, s2,
                           o);
              kin = 0.5 * o[0];
              turning = turn(o, 0) || turn(o, 1);
              break;
            }
            default: {
              double o[7];
              leaf_pass<3>(cx, rs, gs, eps, gi, first, ckr_w, cks_w, r0, s0, r1, s1, r2, s2,
                           o);
              kin = 0.5 * o[0];
              turning = turn(o, 0) || turn(o, 1) || turn(o, 2);
              break;
            }
          }
          next_check = idx_max - nf;
        }
        for (int i = next_check; i >= idx_min && !turning; --i)
          turning = is_turning(cx, cx.v(CKR + i), rnew, srs, cx.v(CKS + i), cx.v(CKR + i));
        const double pe_new = pe_in;
        const double e_new = pe_new + kin;
        const double delta = nan_to_inf(e_new - S.energy_current);
        const double leaf_w = -delta;
        const int leaf_div = delta > cfg.max_delta_energy;
        const double acc = clip_exp_neg(delta);
        float* szp = cx.v(V_SZP);
        float* sgp = cx.v(V_SGP);
        if (leaf_idx == 0) {
          FOR_D(d) { szp[d] = znew[d]; sgp[d] = gnew[d]; }
          S.s_peP = pe_new; S.s_energyP = e_new; S.s_weight = leaf_w;
          S.s_diverging = leaf_div; S.s_sum_accept = acc; S.s_num = 1;
        } else {  // _combine_tree(..., biased_transition=False) (hmc_util.py:767-848)
          const double x = leaf_w - S.s_weight;
          const double e = exp(-fabs(x));
          const double p = expit_e(x, e);
          if (bernoulli(K(S.leaf_key), p, part)) {
            FOR_D(d) { szp[d] = znew[d]; sgp[d] = gnew[d]; }
            S.s_peP = pe_new; S.s_energyP = e_new;
          }
          S.s_weight = log_add_exp_e(S.s_weight, leaf_w, e);
          S.s_diverging = leaf_div;
          S.s_sum_accept += acc;
          S.s_num += 1;
        }
        S.s_turning = turning ? 1 : 0;
        if (S.s_num < (1 << S.depth) && !turning && !S.s_diverging) S.phase = PH_LEAF;
        else S.phase = PH_COMBINE;
        break;
      }
      case PH_COMBINE: {  // _double_tree's _combine_tree(..., True) (hmc_util.py:936-938)
        float* rsum = cx.v(V_RSUM);
        const float* srs = cx.v(V_SRSUM);
        FOR_D(d) rsum[d] += srs[d];
        // min(1, exp(x)) and log_add_exp's exp(-|x|) from one exp: for x <= 0
        // they are the same number, for x > 0 (or NaN) the probability is 1
        const double x = S.s_weight - S.weight;
        const double e = exp(-fabs(x));
        double p = 0.0;
        if (!S.s_turning && !S.s_diverging) p = x <= 0.0 ? e : 1.0;
        const bool turn = S.s_turning || is_turning(cx, cx.v(V_RL), cx.v(V_RR), rsum, nullptr, nullptr);
        if (bernoulli(K(S.top_key), p, part)) {
          vcopy(cx, V_ZP, V_SZP);
          vcopy(cx, V_GP, V_SGP);
          S.peP = S.s_peP;
          S.energyP = S.s_energyP;
        }
        S.depth += 1;
        S.weight = log_add_exp_e(S.weight, S.s_weight, e);
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
        S.pe_edge = pe_in;
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
        bool fs = false;
        if (it < cfg.num_warmup) fs = adapt_update(cfg, cx, S, it, ap);
        if (fs) fs_begin(S, PH_ITER_FINISH);
        else S.phase = PH_ITER_FINISH;
        break;
      }
      case PH_ITER_FINISH: {
        const double ap = S.accept_prob;
        const int it = S.i;
        const int itr = it + 1;
        const int n = (it < cfg.num_warmup) ? itr : itr - cfg.num_warmup;
        S.mean_accept_prob = S.mean_accept_prob + (ap - S.mean_accept_prob) / n;
        S.i = itr;
        collect(cfg, col, cx, S, c, it);
        S.phase = PH_ITER_START;
        break;
      }
      default:
        return need_grad;
    }
  }
  return need_grad;
}

// Vector slots of the position a chain waiting for a gradient asked about and
// of the gradient its waiting phase takes: the leaf's edge, or the right edge
// for the step-size search and fixed-length HMC.
__device__ __forceinline__ void request_slots(const ChainState& S, int& zs, int& gs) {
  const bool left = S.phase == PH_LEAF_WAIT && !S.going_right;
  zs = left ? V_ZL : V_ZR;
  gs = left ? V_GL : V_GR;
}

// Chain index of the calling wavefront (`wpb` wavefronts per block), made
// provably wave-uniform so that per-chain base addresses sit in scalar
// registers instead of occupying a vector register pair each.
__device__ __forceinline__ int wave_chain(int wpb) {
  return __builtin_amdgcn_readfirstlane((int)(blockIdx.x * wpb + threadIdx.x / NAX_WAVE));
}

// ChainState copies by all lanes of a wave.  The kernels keep the state of
// their chains in LDS while they advance them: as a register copy the ~350
// bytes of wave-uniform fields would occupy vector registers throughout the
// state machine and cap the occupancy at one wave per SIMD.
__device__ __forceinline__ void copy_state(ChainState& dst, const ChainState& src, int lane) {
  static_assert(sizeof(ChainState) % 4 == 0, "ChainState must be whole dwords");
  constexpr int n = sizeof(ChainState) / 4;
  const uint32_t* s = reinterpret_cast<const uint32_t*>(&src);
  uint32_t* d = reinterpret_cast<uint32_t*>(&dst);
  for (int i = lane; i < n; i += NAX_WAVE) d[i] = s[i];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DENSE = false gives the diagonal-mass kernels a context whose dense
// pointers are compile-time null, so the dense branches fold away.
// WPC > 1 (a workgroup per chain) is for diagonal mass only: the dense
// factorisation runs on one wavefront.
template <bool DENSE = true, bool LDSV = false, int WPC = 1>
__device__ __forceinline__ ChainCtxT<LDSV, WPC> make_ctx(const EngineCfg& cfg, int c,
                                                         float* vec, double* welf) {
  static_assert(WPC == 1 || !DENSE, "a multi-wave chain group needs the diagonal mass");
  ChainCtxT<LDSV, WPC> cx;
  cx.red = nullptr;
  cx.redf = nullptr;
  cx.lane = WPC == 1 ? (int)(threadIdx.x & (NAX_WAVE - 1)) : (int)threadIdx.x;
  cx.D = cfg.chain_dim ? cfg.chain_dim[c] : cfg.D;
  cx.Dp = cfg.Dp;
  cx.vb = vec + (size_t)c * cfg.nvec * cfg.Dp;
  cx.lb = nullptr;
  cx.wel = welf + (size_t)c * 3 * cfg.Dp;
  const size_t blk = (size_t)c * cfg.Dp * cfg.Dp;
  cx.mi = (DENSE && cfg.dense) ? cfg.dimm_t + blk : nullptr;
  cx.ms = (DENSE && cfg.dense) ? cfg.dmsq_t + blk : nullptr;
  cx.m2 = (DENSE && cfg.dense) ? cfg.dm2 + blk : nullptr;
  return cx;
}

// ---- per-chain potentials ------------------------------------------------------
// A per-chain model M provides
//   template <class G> __device__ double value_and_grad(const G& grp,
//                                                       const float* z, float* g) const
// run by the chain's group G (one wavefront, WaveGroup, or the WPC waves of a
// chain context; grp.lane = thread index in the group, G::NT threads, the
// chain_sum / chain_inclusive_scan primitives): U(z) (returned on every
// thread) and dU/dz written to g.  z and g may be written by other threads of
// the group just before the call, so callers synchronise around it.

template <class M>
__global__ void __launch_bounds__(256)
perchain_eval_kernel(M m, const float* __restrict__ z, int ldz, int C, double* __restrict__ pe,
                     float* __restrict__ g, int ldg) {
  const int c = blockIdx.x * (blockDim.x / NAX_WAVE) + threadIdx.x / NAX_WAVE;
  if (c >= C) return;
  WaveGroup grp;
  grp.lane = threadIdx.x & (NAX_WAVE - 1);
  const double u = m.value_and_grad(grp, z + (size_t)c * ldz, g + (size_t)c * ldg);
  if (grp.lane == 0) pe[c] = u;
}

// The same evaluation with WPC wavefronts (one workgroup) per chain.
template <class M, int WPC>
__global__ void __launch_bounds__(64 * WPC)
perchain_eval_group_kernel(M m, const float* __restrict__ z, int ldz, int C,
                           double* __restrict__ pe, float* __restrict__ g, int ldg) {
  __shared__ double red[kRedSlots * WPC];
  __shared__ float redf[NAX_WAVE * WPC];
  const int c = blockIdx.x;
  if (c >= C) return;
  ChainCtxT<false, WPC> grp;
  grp.red = red;
  grp.redf = redf;
  grp.lane = threadIdx.x;
  const double u = m.value_and_grad(grp, z + (size_t)c * ldz, g + (size_t)c * ldg);
  if (grp.lane == 0) pe[c] = u;
}

// One-shot evaluation of C chains at wpc wavefronts per chain (kGroupSizes):
// one wave per chain packs four chains into a workgroup, more waves give each
// chain a workgroup of its own.
template <class M>
hipError_t launch_group_eval(int wpc, hipStream_t s, const M& m, const float* z, int ldz, int C,
                             double* pe, float* g, int ldg) {
  switch (wpc) {
    case 1:
      hipLaunchKernelGGL(perchain_eval_kernel<M>, dim3((C + 3) / 4), dim3(256), 0, s, m, z, ldz,
                         C, pe, g, ldg);
      break;
#define NAX_GROUP_EVAL(W)                                                                  \
  case W:                                                                                  \
    hipLaunchKernelGGL((perchain_eval_group_kernel<M, W>), dim3(C), dim3(64 * W), 0, s, m, z, \
                       ldz, C, pe, g, ldg);                                                \
    break;
      NAX_GROUP_EVAL(2)
      NAX_GROUP_EVAL(4)
      NAX_GROUP_EVAL(8)
      NAX_GROUP_EVAL(16)
#undef NAX_GROUP_EVAL
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Vector slots a persistent launch of one chain per workgroup keeps in LDS:
// the V_HOT slots every leaf touches when they fit in the CU's 160 KB next to
// the kernel's static LDS, else none.  The host passes nl * Dp * 4 bytes of
// dynamic LDS.
inline int persistent_lds_slots(int Dp, int wpc = 1) {
  // static LDS of the kernel: a ChainState copy per wave + the group's
  // reduction slots, and a margin
  const int budget = 160 * 1024 - 2048 -
                     wpc * (int)(sizeof(ChainState) + kRedSlots * 8 + NAX_WAVE * 4);
  return budget / (Dp * 4) >= V_HOT ? V_HOT : 0;
}

// The whole sampling loop of one chain in one launch: the wave alternates
// advance_chain() and its own potential until the chain reaches `target`.
// WPB wavefronts (chains) per workgroup: one, for latency-bound chains with
// long vectors, spreads the chains over as many CUs (and their L1 caches and
// address units) as there are chains.  LDSV (with WPB = 1): the nl = V_HOT
// leading vector slots live in LDS for the launch (copied in and back out),
// so a leaf's vector passes run at LDS rather than L2 latency; for models
// with long latent vectors.
// WPC > 1 (with WPB = 1): WPC wavefronts of the workgroup form the chain's
// group; each wave keeps its own copy of the chain state, wave 0 writes it
// back.
template <class M, bool DENSE, int WPB, bool LDSV = false, int WPC = 1>
__global__ void __launch_bounds__(64 * WPB * WPC)
persistent_kernel(M m, EngineCfg cfg, Collect col, const int64_t* __restrict__ d_target,
                  ChainState* __restrict__ states, float* __restrict__ vec,
                  double* __restrict__ welf, float* __restrict__ zreq, float* __restrict__ gout,
                  double* __restrict__ peout, int64_t* __restrict__ grad_evals, int nl) {
  static_assert(WPC == 1 || WPB == 1, "a chain group fills its workgroup");
  // the iteration target sits in device memory, so that one captured graph
  // (target step + this launch) replays for every chunk of iterations
  const int64_t target = *d_target;
  __shared__ ChainState sS[WPB * WPC];
  __shared__ double nax_red[WPC > 1 ? kRedSlots * WPC : 1];
  __shared__ float nax_redf[WPC > 1 ? NAX_WAVE * WPC : 1];
  extern __shared__ float nax_slots_lds[];
  const int c = WPC == 1 ? wave_chain(WPB) : (int)blockIdx.x;
  if (c >= cfg.C) return;
  ChainCtxT<LDSV, WPC> cx = make_ctx<DENSE, LDSV, WPC>(cfg, c, vec, welf);
  cx.red = nax_red;
  cx.redf = nax_redf;
  static_assert(!LDSV || WPB == 1, "LDS-resident slots need one chain per workgroup");
  if constexpr (LDSV) {
    const int n = nl * cfg.Dp;
    for (int i = cx.lane; i < n; i += cx.NT) nax_slots_lds[i] = cx.vb[i];
    __syncthreads();
    cx.lb = nax_slots_lds;
  }
  // the first call may take a gradient the host evaluated (initialisation);
  // after that the potential reads the requested position from the chain's
  // own edge slot and writes the gradient next to it, in LDS when the slots
  // live there
  const float* gi = gout + (size_t)c * cfg.Dp;
  const int wave_lane = threadIdx.x & (NAX_WAVE - 1);
  ChainState& S = sS[threadIdx.x / NAX_WAVE];
  copy_state(S, states[c], wave_lane);
  double pe = peout[c];
  int evals = 0;
  while (advance_chain(cfg, col, target, S, cx, c, nullptr, gi, pe, evals)) {
    int zs, gs;
    request_slots(S, zs, gs);
    chain_sync(cx);
    pe = m.value_and_grad(cx, cx.v(zs), cx.v(gs));
    chain_sync(cx);
    gi = cx.v(gs);
  }
  if constexpr (LDSV) {
    __syncthreads();
    const int n = nl * cfg.Dp;
    for (int i = cx.lane; i < n; i += cx.NT) cx.vb[i] = nax_slots_lds[i];
  }
  if (cx.lane < NAX_WAVE) copy_state(states[c], S, wave_lane);
  if (cx.lane == 0) {
    peout[c] = pe;
    if (evals && grad_evals) atomicAdd((unsigned long long*)grad_evals, (unsigned long long)evals);
  }
}


// The same loop for chains whose vectors fit one wavefront (d <= 64) on a
// group of WPC waves: the group's wave 0 alone runs advance_chain() (vector
// passes of one element per lane and wave-level sums, no workgroup barrier),
// the whole group evaluates the potential (rows split over its 64 WPC
// threads).  The head publishes whether it needs a gradient and the request's
// slots in LDS; every wave leaves the loop after the same barrier.
template <class M, bool LDSV, int WPC>
__global__ void __launch_bounds__(64 * WPC)
persistent_head_kernel(M m, EngineCfg cfg, Collect col, const int64_t* __restrict__ d_target,
                       ChainState* __restrict__ states, float* __restrict__ vec,
                       double* __restrict__ welf, float* __restrict__ zreq,
                       float* __restrict__ gout, double* __restrict__ peout,
                       int64_t* __restrict__ grad_evals, int nl) {
  static_assert(WPC > 1, "the head wave form is for chain groups");
  const int64_t target = *d_target;
  __shared__ ChainState sS;
  __shared__ double nax_red[kRedSlots * WPC];
  __shared__ float nax_redf[NAX_WAVE * WPC];
  __shared__ int req[3];  // needs a gradient, position slot, gradient slot
  extern __shared__ float nax_slots_lds[];
  const int c = (int)blockIdx.x;
  if (c >= cfg.C) return;
  const bool head = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) == 0;
  ChainCtxT<LDSV, WPC> gx = make_ctx<false, LDSV, WPC>(cfg, c, vec, welf);
  gx.red = nax_red;
  gx.redf = nax_redf;
  ChainCtxT<LDSV, 1> hx = make_ctx<false, LDSV, 1>(cfg, c, vec, welf);
  if constexpr (LDSV) {
    const int n = nl * cfg.Dp;
    for (int i = gx.lane; i < n; i += gx.NT) nax_slots_lds[i] = gx.vb[i];
    __syncthreads();
    gx.lb = nax_slots_lds;
    hx.lb = nax_slots_lds;
  }
  // the first call may take a gradient the host evaluated (initialisation)
  const float* gi = gout + (size_t)c * cfg.Dp;
  if (head) copy_state(sS, states[c], hx.lane);
  double pe = peout[c];
  int evals = 0;
  for (;;) {
    if (head) {
      const bool more = advance_chain(cfg, col, target, sS, hx, c, nullptr, gi, pe, evals);
      if (hx.lane == 0) {
        int zs, gs;
        request_slots(sS, zs, gs);
        req[0] = more ? 1 : 0;
        req[1] = zs;
        req[2] = gs;
      }
    }
    __syncthreads();
    const int more = req[0], zs = req[1], gs = req[2];
    if (!more) break;
    pe = m.value_and_grad(gx, gx.v(zs), gx.v(gs));
    __syncthreads();  // the gradient is written, req has been read
    gi = gx.v(gs);
  }
  if constexpr (LDSV) {
    const int n = nl * cfg.Dp;
    for (int i = gx.lane; i < n; i += gx.NT) gx.vb[i] = nax_slots_lds[i];
  }
  if (head) {
    copy_state(states[c], sS, hx.lane);
    if (hx.lane == 0) {
      peout[c] = pe;
      if (evals && grad_evals) atomicAdd((unsigned long long*)grad_evals, (unsigned long long)evals);
    }
  }
}

}  // namespace nax

// ---- host: persistent launches with a run-time number of waves per chain ---------
namespace nax {

// Waves per chain the persistent launchers instantiate.
constexpr int kGroupSizes[] = {1, 2, 4, 8, 16};

// Launch shapes tuned so far in this process, keyed by the model's shape:
// every sampler of one shape then runs the same launch (and the same draws
// for the same keys), and pays for the tuning once.
inline bool tuned_shape(const std::string& key, int* value, bool store) {
  static std::mutex mu;
  static std::map<std::string, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  if (store) {
    cache[key] = *value;
    return true;
  }
  auto it = cache.find(key);
  if (it == cache.end()) return false;
  *value = it->second;
  return true;
}

// Advances the device-resident iteration target of persistent launches:
// *t = min(final, *t + step) (one thread).
static __global__ void step_target_kernel(int64_t* t, int64_t final, int64_t step) {
  const int64_t v = *t + step;
  *t = v < final ? v : final;
}
static __global__ void set_target_kernel(int64_t* t, int64_t v) { *t = v; }

// Launches persistent_kernel<M, ., 1, LDSV, wpc> (one chain per workgroup),
// with the hot vector slots in LDS when LDSV and they fit (else the global
// form).  Dense mass runs on one wave.  HEAD (models whose vectors fit one
// wave): groups of more than one wave run persistent_head_kernel instead.
template <class M, bool LDSV, bool HEAD = false>
hipError_t launch_chain_groups(int wpc, hipStream_t s, const M& m, const EngineCfg& cfg,
                               const Collect& col, const int64_t* d_target, ChainState* st,
                               float* vec, double* wel, float* zreq, float* gout,
                               double* peout, int64_t* gevals) {
  if (cfg.dense) wpc = 1;
#define NAX_GROUP_LAUNCH(DN, W)                                                             \
  {                                                                                         \
    if constexpr (HEAD && W > 1 && !DN) {                                                   \
      const int nlh = LDSV ? persistent_lds_slots(cfg.Dp, W) : 0;                           \
      if constexpr (LDSV) {                                                                 \
        if (nlh > 0) {                                                                      \
          const size_t lds = (size_t)nlh * cfg.Dp * sizeof(float);                          \
          static size_t lds_set = 0; /* set once per size (also outside graph capture) */  \
          if (lds > lds_set) {                                                              \
            hipError_t e = hipFuncSetAttribute(                                             \
                (const void*)persistent_head_kernel<M, true, W>,                            \
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                      \
            if (e != hipSuccess) return e;                                                  \
            lds_set = lds;                                                                  \
          }                                                                                 \
          hipLaunchKernelGGL((persistent_head_kernel<M, true, W>), dim3(cfg.C),             \
                             dim3(64 * W), lds, s, m, cfg, col, d_target, st, vec, wel,     \
                             zreq, gout, peout, gevals, nlh);                               \
          return hipGetLastError();                                                         \
        }                                                                                   \
      }                                                                                     \
      hipLaunchKernelGGL((persistent_head_kernel<M, false, W>), dim3(cfg.C), dim3(64 * W),  \
                         0, s, m, cfg, col, d_target, st, vec, wel, zreq, gout, peout,      \
                         gevals, 0);                                                        \
      return hipGetLastError();                                                             \
    }                                                                                       \
    const int nl = LDSV ? persistent_lds_slots(cfg.Dp, W) : 0;                              \
    if constexpr (LDSV) {                                                                   \
      if (nl > 0) {                                                                         \
        const size_t lds = (size_t)nl * cfg.Dp * sizeof(float);                             \
        static size_t lds_set = 0; /* set once per size (also outside graph capture) */    \
        if (lds > lds_set) {                                                                \
          hipError_t e = hipFuncSetAttribute(                                               \
              (const void*)persistent_kernel<M, DN, 1, true, W>,                            \
              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                        \
          if (e != hipSuccess) return e;                                                    \
          lds_set = lds;                                                                    \
        }                                                                                   \
        hipLaunchKernelGGL((persistent_kernel<M, DN, 1, true, W>), dim3(cfg.C),             \
                           dim3(64 * W), lds, s, m, cfg, col, d_target, st, vec, wel, zreq, \
                           gout, peout, gevals, nl);                                        \
        return hipGetLastError();                                                           \
      }                                                                                     \
    }                                                                                       \
    hipLaunchKernelGGL((persistent_kernel<M, DN, 1, false, W>), dim3(cfg.C), dim3(64 * W), \
                       0, s, m, cfg, col, d_target, st, vec, wel, zreq, gout, peout,        \
                       gevals, 0);                                                          \
    return hipGetLastError();                                                               \
  }
  if (cfg.dense) NAX_GROUP_LAUNCH(true, 1)
  switch (wpc) {
    case 1: NAX_GROUP_LAUNCH(false, 1)
    case 2: NAX_GROUP_LAUNCH(false, 2)
    case 4: NAX_GROUP_LAUNCH(false, 4)
    case 8: NAX_GROUP_LAUNCH(false, 8)
    case 16: NAX_GROUP_LAUNCH(false, 16)
    default: return hipErrorInvalidValue;
  }
#undef NAX_GROUP_LAUNCH
}

// Autotuner of a persistent launch parameter: from a scratch copy of the
// current sampler state, each candidate runs the real persistent kernel
// (launch(cand, stream, Collect, d_target, st, vec, wel, zreq, gout, peout,
// gevals)) for `iters` iterations -- the same NUTS work, nothing collected --
// timed with HIP events; the fastest candidate is returned.  The sampler's
// own state is not touched.  `log` receives "candidate:ms" for each (the
// candidate written by `label` when given).
template <class Launch>
int tune_persistent(hipStream_t s, const EngineCfg& cfg, const ChainState* st, const float* vec,
                    const double* wel, const float* zreq, const float* gout, const double* peout,
                    int iters, const std::vector<int>& cands, Launch launch, std::string* log,
                    std::string (*label)(int) = nullptr) {
  if (cands.size() < 2) return cands.empty() ? 1 : cands[0];
  const size_t nst = sizeof(ChainState) * cfg.C;
  const size_t nvec = sizeof(float) * (size_t)cfg.C * cfg.nvec * cfg.Dp;
  const size_t nwel = sizeof(double) * (size_t)cfg.C * 3 * cfg.Dp;
  const size_t nzg = sizeof(float) * (size_t)cfg.C * cfg.Dp;
  const size_t npe = sizeof(double) * cfg.C;
  ChainState* st2 = nullptr;
  float *vec2 = nullptr, *zr2 = nullptr, *g2 = nullptr;
  double *wel2 = nullptr, *pe2 = nullptr;
  int64_t* ge2 = nullptr;  // [0]: grad evals, [1]: the target
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<ChainState> hs(cfg.C);
  bool ok = hipMalloc(&st2, nst) == hipSuccess && hipMalloc(&vec2, nvec) == hipSuccess &&
            hipMalloc(&wel2, nwel) == hipSuccess && hipMalloc(&zr2, nzg) == hipSuccess &&
            hipMalloc(&g2, nzg) == hipSuccess && hipMalloc(&pe2, npe) == hipSuccess &&
            hipMalloc(&ge2, 2 * sizeof(int64_t)) == hipSuccess &&
            hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
            hipMemcpyAsync(hs.data(), st, nst, hipMemcpyDeviceToHost, s) == hipSuccess &&
            hipStreamSynchronize(s) == hipSuccess;
  int best = cands[0];
  if (ok) {
    int64_t i0 = 0;
    for (const ChainState& c : hs) i0 = std::max<int64_t>(i0, c.i);
    Collect none;
    memset(&none, 0, sizeof(none));
    float best_ms = 1e30f;
    for (int w : cands) {
      float ms = 1e30f;
      hipLaunchKernelGGL(set_target_kernel, dim3(1), dim3(1), 0, s, ge2 + 1, i0 + iters);
      if (hipGetLastError() == hipSuccess &&
          hipMemcpyAsync(st2, st, nst, hipMemcpyDeviceToDevice, s) == hipSuccess &&
          hipMemcpyAsync(vec2, vec, nvec, hipMemcpyDeviceToDevice, s) == hipSuccess &&
          hipMemcpyAsync(wel2, wel, nwel, hipMemcpyDeviceToDevice, s) == hipSuccess &&
          hipMemcpyAsync(zr2, zreq, nzg, hipMemcpyDeviceToDevice, s) == hipSuccess &&
          hipMemcpyAsync(g2, gout, nzg, hipMemcpyDeviceToDevice, s) == hipSuccess &&
          hipMemcpyAsync(pe2, peout, npe, hipMemcpyDeviceToDevice, s) == hipSuccess &&
          hipEventRecord(e0, s) == hipSuccess &&
          launch(w, s, none, ge2 + 1, st2, vec2, wel2, zr2, g2, pe2, ge2) == hipSuccess &&
          hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess)
        hipEventElapsedTime(&ms, e0, e1);
      if (log) *log += " " + (label ? label(w) : std::to_string(w)) + ":" + std::to_string(ms);
      if (ms < best_ms) {
        best_ms = ms;
        best = w;
      }
    }
  }
  hipFree(st2); hipFree(vec2); hipFree(wel2); hipFree(zr2); hipFree(g2); hipFree(pe2);
  hipFree(ge2);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  (void)hipGetLastError();  // a failed candidate leaves no sticky error behind
  return best;
}

// Waves per chain of a persistent per-chain model (kGroupSizes, one wave
// under dense mass), by tune_persistent.
template <class M, bool LDSV, bool HEAD = false>
int tune_chain_groups(hipStream_t s, const M& m, const EngineCfg& cfg, const ChainState* st,
                      const float* vec, const double* wel, const float* zreq, const float* gout,
                      const double* peout, int iters, std::string* log) {
  if (cfg.dense) return 1;
  std::vector<int> cands(std::begin(kGroupSizes), std::end(kGroupSizes));
  auto launch = [&](int w, hipStream_t ss, const Collect& col, const int64_t* tgt,
                    ChainState* st2, float* vec2, double* wel2, float* zr2, float* g2,
                    double* pe2, int64_t* ge2) {
    return launch_chain_groups<M, LDSV, HEAD>(w, ss, m, cfg, col, tgt, st2, vec2, wel2, zr2, g2,
                                              pe2, ge2);
  };
  return tune_persistent(s, cfg, st, vec, wel, zreq, gout, peout, iters, cands, launch, log);
}

}  // namespace nax
