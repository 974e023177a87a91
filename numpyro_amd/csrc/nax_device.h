// Device helpers shared by the engine kernels: threefry2x32-20 and the
// jax.random key/bit/float derivations, erfinv, wave reductions.
//
// Random streams follow jax.random as numpyro 0.21.0 consumes it:
//   split:   numpyro/infer/hmc.py:335,472  numpyro/infer/hmc_util.py:565,656,920,1005,1161
//   normal:  numpyro/infer/hmc.py:102
//   bernoulli: numpyro/infer/hmc_util.py:804,1162
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NAX_WAVE 64

namespace nax {

struct Key {
  uint32_t k0, k1;
};

__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) {
  return (x << r) | (x >> (32 - r));
}

// threefry2x32 with 20 rounds: rotations {13,15,26,6} / {17,29,16,24},
// key injection after every 4 rounds (Salmon et al., SC'11).
__device__ __forceinline__ void threefry2x32(Key k, uint32_t& x0, uint32_t& x1) {
  const uint32_t ks0 = k.k0, ks1 = k.k1, ks2 = k.k0 ^ k.k1 ^ 0x1BD11BDAu;
  x0 += ks0;
  x1 += ks1;
#define NAX_R(r) { x0 += x1; x1 = rotl32(x1, r); x1 ^= x0; }
  NAX_R(13) NAX_R(15) NAX_R(26) NAX_R(6)
  x0 += ks1; x1 += ks2 + 1u;
  NAX_R(17) NAX_R(29) NAX_R(16) NAX_R(24)
  x0 += ks2; x1 += ks0 + 2u;
  NAX_R(13) NAX_R(15) NAX_R(26) NAX_R(6)
  x0 += ks0; x1 += ks1 + 3u;
  NAX_R(17) NAX_R(29) NAX_R(16) NAX_R(24)
  x0 += ks1; x1 += ks2 + 4u;
  NAX_R(13) NAX_R(15) NAX_R(26) NAX_R(6)
  x0 += ks2; x1 += ks0 + 5u;
#undef NAX_R
}

// jax "threefry_2x32(key, iota(n))" word i (original derivation): the count
// array [0..n) (padded to even length) is cut into halves that form the two
// cipher inputs; output = concat(out0, out1).
__device__ __forceinline__ uint32_t tf_iota_word(Key k, uint32_t i, uint32_t n) {
  const uint32_t h = (n + 1) / 2;
  const uint32_t j = (i < h) ? i : i - h;
  uint32_t x0 = j, x1 = (j + h < n) ? j + h : 0u;
  threefry2x32(k, x0, x1);
  return (i < h) ? x0 : x1;
}

// random.split(key, num)[idx]
__device__ __forceinline__ Key split_key(Key k, uint32_t idx, uint32_t num, bool part) {
  if (part) {
    uint32_t x0 = 0, x1 = idx;
    threefry2x32(k, x0, x1);
    return Key{x0, x1};
  }
  return Key{tf_iota_word(k, 2 * idx, 2 * num), tf_iota_word(k, 2 * idx + 1, 2 * num)};
}

// 32-bit random_bits(key, (n,))[i]
__device__ __forceinline__ uint32_t bits_vec(Key k, uint32_t i, uint32_t n, bool part) {
  if (part) {
    uint32_t x0 = 0, x1 = i;
    threefry2x32(k, x0, x1);
    return x0 ^ x1;
  }
  return tf_iota_word(k, i, n);
}

// 32-bit random_bits(key, ()) (scalar shape)
__device__ __forceinline__ uint32_t bits_scalar(Key k, bool part) {
  uint32_t x0 = 0, x1 = 0;
  threefry2x32(k, x0, x1);
  return part ? (x0 ^ x1) : x0;
}

// [0,1) float32 from 32 random bits: (bits >> 9) | 0x3F800000, minus 1.
__device__ __forceinline__ float bits_to_unit(uint32_t b) {
  return __uint_as_float((b >> 9) | 0x3F800000u) - 1.0f;
}

// Inverse error function, single precision, from M. Giles, "Approximating
// the erfinv function" (GPU Computing Gems Jade, 2011): central branch for
// w = -log(1-x^2) < 5, tail branch in sqrt(w) otherwise.
__device__ __forceinline__ float erfinv_giles(float x) {
  float w = -log1pf(-x * x);
  float p;
  if (w < 5.0f) {
    w = w - 2.5f;
    p = 2.81022636e-08f;
    p = fmaf(p, w, 3.43273939e-07f);
    p = fmaf(p, w, -3.5233877e-06f);
    p = fmaf(p, w, -4.39150654e-06f);
    p = fmaf(p, w, 0.00021858087f);
    p = fmaf(p, w, -0.00125372503f);
    p = fmaf(p, w, -0.00417768164f);
    p = fmaf(p, w, 0.246640727f);
    p = fmaf(p, w, 1.50140941f);
  } else {
    w = sqrtf(w) - 3.0f;
    p = -0.000200214257f;
    p = fmaf(p, w, 0.000100950558f);
    p = fmaf(p, w, 0.00134934322f);
    p = fmaf(p, w, -0.00367342844f);
    p = fmaf(p, w, 0.00573950773f);
    p = fmaf(p, w, -0.0076224613f);
    p = fmaf(p, w, 0.00943887047f);
    p = fmaf(p, w, 1.00167406f);
    p = fmaf(p, w, 2.83297682f);
  }
  return p * x;
}

// random.normal(key, (n,))[i]: u ~ U(nextafter(-1, 0), 1) in float32
// (f * 2 + lo, clamped at lo), then sqrt(2) * erfinv(u).
__device__ __forceinline__ float normal_vec(Key k, uint32_t i, uint32_t n, bool part) {
  const float lo = -0.99999994f;  // nextafter(-1, 0) in float32
  float f = bits_to_unit(bits_vec(k, i, n, part));
  float u = fmaxf(lo, f * 2.0f + lo);
  return 1.41421356237309515f * erfinv_giles(u);
}

// random.uniform(key, (n,), minval, maxval)[i]
__device__ __forceinline__ float uniform_vec(Key k, uint32_t i, uint32_t n, bool part,
                                             float lo, float hi) {
  float f = bits_to_unit(bits_vec(k, i, n, part));
  return fmaxf(lo, f * (hi - lo) + lo);
}

// Positive latent sites in unconstrained space (potential_energy's
// _unconstrain_reparam, numpyro/infer/util.py:300-330): biject_to(positive)
// is ExpTransform (transforms.py:606-652, 2115-2118; x = e^u, log|J| = u),
// biject_to(softplus_positive) is SoftplusTransform (transforms.py:1156-1175,
// 2213-2215; x = softplus(u) = log1p(e^u), log|J| = -softplus(-u)).  For u:
// x, dx/du, log|dx/du| and its derivative d log|dx/du| / du.
struct PosTf {
  double x, dx, ladj, dladj;
};
__device__ __forceinline__ PosTf positive_tf(double u, bool softplus) {
  if (!softplus) {
    const double x = exp(u);
    return PosTf{x, x, u, 1.0};
  }
  const double e = exp(-fabs(u));
  const double sp = fmax(u, 0.0) + log1p(e);            // softplus(u), no overflow
  const double sg = u >= 0.0 ? 1.0 / (1.0 + e) : e / (1.0 + e);   // sigmoid(u)
  const double sgn = u >= 0.0 ? e / (1.0 + e) : 1.0 / (1.0 + e);  // sigmoid(-u)
  return PosTf{sp, sg, u - sp, sgn};                    // -softplus(-u) = u - softplus(u)
}

// Neumaier's compensated summation (improved Kahan-Babuska, Neumaier 1974):
// `s + c` carries the running sum with the rounding error of every addition
// kept in `c`.  Used for the fp64 energy sums across tiles and partials.
struct CompSum {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ void add(const CompSum& o) {
    add(o.s);
    c += o.c;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

// random.bernoulli(key, p) for a scalar p: uniform(key, ()) < p.
__device__ __forceinline__ bool bernoulli(Key k, double p, bool part) {
  return (double)bits_to_unit(bits_scalar(k, part)) < p;
}

// ---- wave (64-lane) reductions ---------------------------------------------
// Called with all 64 lanes active.  Lanes combine inside rows of 16 through
// DPP lane permutations (VALU moves, no LDS round trip as a shuffle takes);
// the row totals are read out of lanes 0/16/32/48 (or 15/31/47 for a scan)
// and combined in lane order, so every lane returns the same value.
namespace dpp {
constexpr int kXor1 = 0xB1;        // quad_perm [1, 0, 3, 2]
constexpr int kXor2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int kHalfMirror = 0x141;  // lane i <-> 7 - i inside each 8 lanes
constexpr int kMirror = 0x140;      // lane i <-> 15 - i inside each row
constexpr int kShr1 = 0x111;        // row_shr: lane i reads i - n of its row (0 below)
constexpr int kShr2 = 0x112;
constexpr int kShr4 = 0x114;
constexpr int kShr8 = 0x118;
}  // namespace dpp

template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF,
                                                            false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL,
                                                            0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float read_lane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <class T>
__device__ __forceinline__ T wave_sum_t(T v) {
  v += dpp_mov<dpp::kXor1>(v);  // pairs
  v += dpp_mov<dpp::kXor2>(v);  // quads
  v += dpp_mov<dpp::kHalfMirror>(v);  // half rows
  v += dpp_mov<dpp::kMirror>(v);  // rows
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}
__device__ __forceinline__ double wave_sum(double v) { return wave_sum_t(v); }
__device__ __forceinline__ float wave_sum(float v) { return wave_sum_t(v); }

// Inclusive prefix sum over the wave's lanes in lane order: a scan inside
// each row of 16, then the totals of the rows before.
__device__ __forceinline__ double wave_inclusive_scan(double v, int lane) {
  v += dpp_mov<dpp::kShr1>(v);
  v += dpp_mov<dpp::kShr2>(v);
  v += dpp_mov<dpp::kShr4>(v);
  v += dpp_mov<dpp::kShr8>(v);
  const double t0 = read_lane(v, 15), t1 = read_lane(v, 31), t2 = read_lane(v, 47);
  const int row = lane >> 4;
  const double off = row == 0 ? 0.0 : row == 1 ? t0 : row == 2 ? t0 + t1 : (t0 + t1) + t2;
  return v + off;
}

// Work split of the MFMA logistic kernels for n_act requesting chains over a
// launch of nblk workgroups: CT tiles of 128 chains times RB row splits (a
// multiple of the 8 XCDs, at most one split per 8 row tiles), so that the
// whole launch stays busy however few chains ask for a gradient.
__host__ __device__ __forceinline__ void logit_split(int n_act, int nblk, int64_t ntile, int& CT,
                                                     int& RB) {
  CT = n_act > 0 ? (n_act + 127) / 128 : 1;
  int rb = (nblk / CT) / 8 * 8;
  rb = rb < 8 ? 8 : rb;
  const int64_t cap = (ntile + 7) / 8 * 8;
  RB = rb < cap ? rb : (int)cap;
}

__device__ __forceinline__ double log_add_exp(double a, double b) {
  double m = fmax(a, b);
  if (m == -INFINITY) return -INFINITY;
  return m + log1p(exp(-fabs(a - b)));
}

__device__ __forceinline__ double expit(double x) {
  return x >= 0 ? 1.0 / (1.0 + exp(-x)) : exp(x) / (1.0 + exp(x));
}

// expit(x) and log_add_exp(a, b) for a - b = -x from e = exp(-|x|), which
// both evaluate: the same values as the two functions, one exp less.
__device__ __forceinline__ double expit_e(double x, double e) {
  return x >= 0 ? 1.0 / (1.0 + e) : e / (1.0 + e);
}
__device__ __forceinline__ double log_add_exp_e(double a, double b, double e) {
  const double m = fmax(a, b);
  return m == -INFINITY ? -INFINITY : m + log1p(e);
}

// Bernoulli-logits terms of one observation: r = sigmoid(l) - y (dU/dl) and
// u = softplus(l) - y l (= -log Bernoulli(y | logits l),
// numpyro/distributions/util.py:317-320), from e = exp(-|l|) with the
// hardware exp2 / rcp / log2 (each within ~1 ulp).
__device__ __forceinline__ void logistic_terms_f(float l, float y, float& r, float& u) {
  const float e = __builtin_amdgcn_exp2f(-fabsf(l) * 1.44269504088896341f);
  const float d = 1.0f + e;
  const float inv = __builtin_amdgcn_rcpf(d);
  const float s = l >= 0.0f ? inv : e * inv;
  r = s - y;
  u = fmaf(__builtin_amdgcn_logf(d), 0.693147180559945309f, fmaxf(l, 0.0f) - y * l);
}

}  // namespace nax
