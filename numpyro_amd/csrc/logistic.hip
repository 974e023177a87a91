// Bayesian logistic regression potential and its hand-written adjoint for C
// chains at once.
//
//   U(b) = sum_n [softplus(l_n) - y_n l_n] + sum_d [b_d^2 / (2 s^2) + log(sqrt(2 pi) s)]
//   dU   = X^T (sigmoid(l) - y) + b / s^2,        l = X b
//
// which is -(BernoulliLogits.log_prob + Normal.log_prob) of the reference
// (numpyro/distributions/discrete.py:244-263, distributions/util.py:317-320,
// distributions/continuous.py:2975-2989) for the model of
// examples/covtype.py:66-71, with jax.value_and_grad (hmc_util.py:242-252)
// replaced by the analytic adjoint.  Over C chains, L = X Z^T and
// G = X^T R are two GEMMs that share each X row tile: a block loads a tile
// of X once and uses it for both.  Row tiles are split over blocks; each
// block writes fp32 gradient partials and fp64 energy partials, and a second
// kernel sums the partials in a fixed order (deterministic) and adds the
// prior.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "nax_device.h"
#include "nax_internal.h"

namespace nax {

__device__ __forceinline__ float softplusf_(float x) {
  // max(x, 0) + log1p(exp(-|x|))  (distributions/util.py:317-320)
  return fmaxf(x, 0.0f) + log1pf(__expf(-fabsf(x)));
}
__device__ __forceinline__ float sigmoidf_(float x) { return 1.0f / (1.0f + __expf(-x)); }

// ---- X storage forms of the tile kernel ----------------------------------------
// fp32: X as given.  FP8: X[n][d] = e4m3(code[n][d]) * scale[d], the OCP
// e4m3fn byte (1 sign, 4 exponent bits with bias 7, 3 mantissa bits; no
// infinities) with a power-of-two scale per column; the value of a code
// comes from a 256-entry table, so the products are exact in fp32.
__constant__ float kE4M3[256];

struct XF32 {
  const float* x;
  __device__ __forceinline__ float operator()(int64_t n, int d, int D) const {
    return x[n * D + d];
  }
};
struct XFP8 {
  const uint8_t* q;
  const float* scale;
  __device__ __forceinline__ float operator()(int64_t n, int d, int D) const {
    return kE4M3[q[n * D + d]] * scale[d];
  }
};

// Round-to-nearest-even e4m3fn code of v with |v| <= 448 (subnormal step
// 2^-9; codes count up through the binades, so a mantissa carry moves to
// the next exponent by itself).
__device__ __forceinline__ uint8_t e4m3_encode(float v) {
  const uint8_t sign = v < 0.0f ? 0x80 : 0;
  const float a = fabsf(v);
  int code;
  if (a < 0.015625f) {                       // below 2^-6: subnormal step 2^-9
    code = (int)rintf(a * 512.0f);
  } else {
    int e;
    frexpf(a, &e);                           // a = m 2^e, m in [0.5, 1)
    e -= 1;                                  // leading bit 2^e, e in [-6, 8]
    const int q = (int)rintf(ldexpf(a, 3 - e));   // 8..16, exact scaling
    code = ((e + 7) << 3) + (q - 8);
  }
  return (uint8_t)(sign | (code > 0x7E ? 0x7E : code));
}

// Per-column scale: the power of two s >= max_n |X[n][d]| / 448 (1 for an
// all-zero column), one workgroup per column.
__global__ void __launch_bounds__(256)
fp8_column_scales(const float* __restrict__ X, int64_t N, int D, float* __restrict__ scale) {
  const int d = blockIdx.x;
  float m = 0.0f;
  for (int64_t n = threadIdx.x; n < N; n += 256) m = fmaxf(m, fabsf(X[n * D + d]));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    float s = 1.0f;
    if (m > 0.0f) {
      int e;
      frexpf(m / 448.0f, &e);                // m / 448 = f 2^e, f in [0.5, 1)
      s = ldexpf(1.0f, e);                   // 2^e >= m / 448
      if (s * 0.5f >= m / 448.0f) s *= 0.5f;  // m / 448 a power of two itself
    }
    scale[d] = s;
  }
}

__global__ void fp8_encode(const float* __restrict__ X, int64_t total, int D,
                           const float* __restrict__ scale, uint8_t* __restrict__ q) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x)
    q[i] = e4m3_encode(X[i] / scale[i % D]);  // division by a power of two: exact
}

// ---- the VALU tile kernel ---------------------------------------------------------
// Block: 256 threads, BN rows x BC chains per step.  DCH = ceil(D/64) column
// chunks held per thread in the gradient accumulator.
template <int BN, int BC, int DCH, class XL = XF32>
__global__ void __launch_bounds__(256)
logit_partial_f32(XL X, const float* __restrict__ y, int64_t N, int D,
                  const float* __restrict__ Z, int ldz, int C, int RB,
                  float* __restrict__ wsg, double* __restrict__ wsu) {
  extern __shared__ float smem[];
  const int DP = DCH * 64 + 1;
  float* Zs = smem;                       // [BC][DP]
  float* Xs = Zs + BC * DP;               // [BN][DP]
  float* Rs = Xs + BN * DP;               // [BN][BC]
  double* Ud = (double*)(Rs + BN * BC);   // [256]
  const int t = threadIdx.x;
  const int c0 = blockIdx.y * BC;
  const int rb = blockIdx.x;

  for (int i = t; i < BC * DP; i += 256) {
    int c = i / DP, d = i % DP;
    Zs[i] = (c0 + c < C && d < D) ? Z[(size_t)(c0 + c) * ldz + d] : 0.0f;
  }
  // gradient accumulators: thread owns d = (t & 63) + 64 j, c = (t >> 6) + 4 k
  constexpr int CK = BC / 4;
  float acc[DCH][CK];
#pragma unroll
  for (int j = 0; j < DCH; ++j)
#pragma unroll
    for (int k = 0; k < CK; ++k) acc[j][k] = 0.0f;
  CompSum uacc;  // per (row, chain) pair owned in the L phase, compensated
  constexpr int PAIRS = BN * BC / 256;
  static_assert(PAIRS >= 1 && (BN * BC) % 256 == 0, "tile");

  const int64_t ntiles = (N + BN - 1) / BN;
  for (int64_t tile = rb; tile < ntiles; tile += RB) {
    const int64_t n0 = tile * BN;
    __syncthreads();
    // X tile: rows n0..n0+BN-1 are contiguous in row-major X
    for (int i = t; i < BN * D; i += 256) {
      int r = i / D, d = i % D;
      int64_t n = n0 + r;
      Xs[r * DP + d] = (n < N) ? X(n, d, D) : 0.0f;
    }
    __syncthreads();
    // L = X_tile Z^T, R = sigmoid(L) - y, U += softplus(L) - y L
#pragma unroll
    for (int p = 0; p < PAIRS; ++p) {
      int q = t + 256 * p;
      int r = q % BN, c = q / BN;
      const float* xr = Xs + r * DP;
      const float* zc = Zs + c * DP;
      float l = 0.0f;
      for (int d = 0; d < D; ++d) l = fmaf(xr[d], zc[d], l);
      int64_t n = n0 + r;
      float rv = 0.0f;
      if (n < N && c0 + c < C) {
        float yv = y[n];
        rv = sigmoidf_(l) - yv;
        uacc.add((double)(softplusf_(l) - yv * l));
      }
      Rs[r * BC + c] = rv;
    }
    __syncthreads();
    // G[c][d] += sum_r X[r][d] R[r][c]
    for (int r = 0; r < BN; ++r) {
      const float* xr = Xs + r * DP;
      const float* rr = Rs + r * BC;
#pragma unroll
      for (int j = 0; j < DCH; ++j) {
        float xv = xr[(t & 63) + 64 * j];
#pragma unroll
        for (int k = 0; k < CK; ++k) acc[j][k] = fmaf(xv, rr[(t >> 6) + 4 * k], acc[j][k]);
      }
    }
  }
  // write gradient partials [rb][c][d]
#pragma unroll
  for (int j = 0; j < DCH; ++j) {
    int d = (t & 63) + 64 * j;
#pragma unroll
    for (int k = 0; k < CK; ++k) {
      int c = c0 + (t >> 6) + 4 * k;
      if (d < D && c < C) wsg[((size_t)rb * C + c) * D + d] = acc[j][k];
    }
  }
  // energy partials: pairs q = t + 256 p have chain q / BN; with PAIRS == 1
  // threads [c*BN, (c+1)*BN) own chain c.
  Ud[t] = uacc.value();
  __syncthreads();
  if (t < BC && c0 + t < C) {
    CompSum s;
    for (int q = t * BN; q < (t + 1) * BN; ++q) s.add(Ud[q % 256]);
    wsu[(size_t)rb * C + c0 + t] = s.value();
  }
}

// Sums the row-split partials and adds the Normal(0, s) prior; block ci
// reduces workspace column ci.  Gradient: with D <= 128 the 256 threads form
// 256 / DP2 groups (DP2 = D rounded up to a power of two), thread (group q,
// column d) adds the splits q, q + groups, ... in order and the group sums
// are added in group order; wider rows take a column per thread.  Energy:
// thread t adds the splits t, t + 256, ... into a compensated sum and the
// 256 sums are combined in a fixed pairwise tree.  Every order is fixed, so
// results repeat bit for bit; the splits are read by all threads at once
// instead of one column's (or, for the energy, one thread's) dependent loop.
__global__ void __launch_bounds__(256)
logit_reduce(const float* __restrict__ wsg, const double* __restrict__ wsu, int RB, int C,
             int D, const float* __restrict__ Z, int ldz, double inv_s2, double log_norm,
             double u_offset, double* __restrict__ pe, float* __restrict__ g, int ldg,
             const int* __restrict__ act, const int* __restrict__ nact, int nblk,
             int64_t ntile) {
  // workspace column ci = blockIdx.x holds chain act[ci] (ci itself when act
  // is null); columns past *nact were not computed.  nblk > 0: the MFMA
  // kernels' split of an nblk-workgroup launch over the requesting chains
  // (RB row splits, row stride CT 128); nblk = 0: RB as given, stride C.
  const int ci = blockIdx.x;
  const int n_act = act ? *nact : C;
  if (ci >= n_act) return;
  const int c = act ? act[ci] : ci;
  int ldw = C;
  if (nblk > 0) {
    int CT;
    logit_split(n_act, nblk, ntile, CT, RB);
    ldw = CT * 128;
  }
  const int t = threadIdx.x;
  __shared__ double red[256];
  __shared__ double us[256], uc[256];
  __shared__ float gred[256];
  int dp2 = 1;
  while (dp2 < D) dp2 <<= 1;
  double pr = 0.0;
  if (dp2 <= 128) {
    const int ng = 256 / dp2, q = t / dp2, d = t % dp2;
    float sg = 0.0f;
    if (d < D)
      for (int r = q; r < RB; r += ng) sg += wsg[((size_t)r * ldw + ci) * D + d];
    gred[t] = sg;
    __syncthreads();
    if (q == 0 && d < D) {
      float tot = 0.0f;
      for (int k = 0; k < ng; ++k) tot += gred[k * dp2 + d];
      const float zv = Z[(size_t)c * ldz + d];
      g[(size_t)c * ldg + d] = tot + (float)(zv * inv_s2);
      pr = 0.5 * (double)zv * zv * inv_s2 + log_norm;
    }
  } else {
    for (int d = t; d < D; d += 256) {
      float sg = 0.0f;
      for (int r = 0; r < RB; ++r) sg += wsg[((size_t)r * ldw + ci) * D + d];
      const float zv = Z[(size_t)c * ldz + d];
      g[(size_t)c * ldg + d] = sg + (float)(zv * inv_s2);
      pr += 0.5 * (double)zv * zv * inv_s2 + log_norm;
    }
  }
  CompSum u;  // this thread's row-split energy partials, compensated
  for (int r = t; r < RB; r += 256) u.add(wsu[(size_t)r * ldw + ci]);
  red[t] = pr;
  us[t] = u.s;
  uc[t] = u.c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) {
      red[t] += red[t + o];
      CompSum a, b;
      a.s = us[t];
      a.c = uc[t];
      b.s = us[t + o];
      b.c = uc[t + o];
      a.add(b);
      us[t] = a.s;
      uc[t] = a.c;
    }
    __syncthreads();
  }
  if (t == 0) {
    CompSum tot;
    tot.s = us[0];
    tot.c = uc[0];
    tot.add(red[0]);
    tot.add(u_offset);
    pe[c] = tot.value();
  }
}

struct LogisticPotential : Potential {
  const float* x = nullptr;
  const float* y = nullptr;
  int64_t n = 0;
  int d = 0;
  double prior_scale = 1.0;
  int x_dtype = NAX_X_F32;
  uint8_t* q8 = nullptr;     // FP8 form: e4m3 codes [n][d]
  float* scale8 = nullptr;   // and the per-column scales [d]
  float* wsg = nullptr;
  double* wsu = nullptr;
  size_t ws_chains = 0;
  int rb = 0;
  int rb_pick = 0;  // row splits chosen by the autotuner (0: default rule)

  ~LogisticPotential() override {
    hipFree(wsg);
    hipFree(wsu);
    hipFree(q8);
    hipFree(scale8);
  }
  int dim() const override { return d; }
  const char* name() const override {
    return x_dtype == NAX_X_FP8 ? "logit_partial_fp8" : "logit_partial_f32";
  }
  double alg_bytes(int C) const override {
    return (x_dtype == NAX_X_FP8 ? 1.0 : 4.0) * (double)n * d + 4.0 * n;
  }
  double alg_flops(int C) const override { return 4.0 * (double)n * d * C; }

  hipError_t ensure_ws(int C, int RB) {
    if (ws_chains >= (size_t)C && rb >= RB) return hipSuccess;
    hipFree(wsg);
    hipFree(wsu);
    wsg = nullptr;
    wsu = nullptr;
    hipError_t e = hipMalloc(&wsg, sizeof(float) * (size_t)RB * C * d);
    if (e != hipSuccess) return e;
    e = hipMalloc(&wsu, sizeof(double) * (size_t)RB * C);
    if (e != hipSuccess) return e;
    ws_chains = C;
    rb = RB;
    return hipSuccess;
  }

  hipError_t eval(hipStream_t s, const float* z, int ldz, int C, double* pe, float* g,
                  int ldg) override {
    constexpr int BN = 32, BC = 8;
    const int ctiles = (C + BC - 1) / BC;
    const int64_t ntiles = (n + BN - 1) / BN;
    int RB = (int)std::min<int64_t>(ntiles, std::max<int64_t>(1, 2048 / ctiles));
    if (rb_pick > 0) RB = (int)std::min<int64_t>(ntiles, rb_pick);
    hipError_t e = ensure_ws(C, RB);
    if (e != hipSuccess) return e;
    const int dch = (d + 63) / 64;
    const int DP = dch * 64 + 1;
    size_t shm = sizeof(float) * ((size_t)BC * DP + (size_t)BN * DP + BN * BC) + 256 * sizeof(double);
    dim3 grid(RB, ctiles);
    mark_begin(s);
#define NAX_LAUNCH(DC)                                                                       \
  if (x_dtype == NAX_X_FP8)                                                                  \
    hipLaunchKernelGGL((logit_partial_f32<BN, BC, DC, XFP8>), grid, dim3(256), shm, s,        \
                       XFP8{q8, scale8}, y, n, d, z, ldz, C, RB, wsg, wsu);                   \
  else                                                                                       \
    hipLaunchKernelGGL((logit_partial_f32<BN, BC, DC, XF32>), grid, dim3(256), shm, s,        \
                       XF32{x}, y, n, d, z, ldz, C, RB, wsg, wsu)
    switch (dch) {
      case 1: NAX_LAUNCH(1); break;
      case 2: NAX_LAUNCH(2); break;
      case 3: NAX_LAUNCH(3); break;
      case 4: NAX_LAUNCH(4); break;
      case 5: NAX_LAUNCH(5); break;
      case 6: NAX_LAUNCH(6); break;
      case 7: NAX_LAUNCH(7); break;
      case 8: NAX_LAUNCH(8); break;
      default: return hipErrorInvalidValue;
    }
#undef NAX_LAUNCH
    mark_end(s);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const double inv_s2 = 1.0 / (prior_scale * prior_scale);
    const double log_norm = std::log(std::sqrt(2.0 * M_PI) * prior_scale);
    hipLaunchKernelGGL(logit_reduce, dim3(C), dim3(256), 0, s, wsg, wsu, RB, C, d, z, ldz,
                       inv_s2, log_norm, 0.0, pe, g, ldg, nullptr, nullptr, 0, 0);
    return hipGetLastError();
  }
};

// ---- the allocation-free entry point ---------------------------------------------
// The row-split partials of logit_partial_f32 / logit_reduce live in this
// file-scope device array instead of a hipMalloc'd workspace: fp64 energy
// partials [RB][C] first, then fp32 gradient partials [RB][C][D].  Each GPU
// holds its own copy until the process ends.  Host code never passes the
// array's name to a kernel: it passes the device address that
// hipGetSymbolAddress returned for the current device.
constexpr size_t kScratchBytes = size_t(16) << 20;  // 16 MiB
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device, looked up once per
// device (with that device current).
static hipError_t scratch_address(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

// Row splits for `count` chains: the VALU tile kernel's rule (as the
// handle-based fp32 tile potential uses it), reduced until the partials fit
// the scratch array; 0 when even one split does not fit.
static static int noalloc_row_splits(int64_t n, int d, int count) {
  constexpr int BN = 32, BC = 8;
  const int ctiles = (count + BC - 1) / BC;
  const int64_t ntiles = (n + BN - 1) / BN;
  int64_t rb = std::min<int64_t>(ntiles, std::max<int64_t>(1, 2048 / ctiles));
  const size_t per = (size_t)count * (sizeof(double) + sizeof(float) * (size_t)d);
  const size_t fit = per ? (kScratchBytes - 256) / per : 0;
  if (fit < 1) return 0;
  return (int)std::min<int64_t>(rb, (int64_t)fit);
}

// The e4m3fn value of every code (OCP FP8: no infinities, S.1111.111 NaN).
static void fill_e4m3_table(float* t) {
  for (int c = 0; c < 256; ++c) {
    const int e = (c >> 3) & 15, m = c & 7;
    float v;
    if (e == 15 && m == 7) v = NAN;
    else if (e == 0) v = std::ldexp((float)m, -9);
    else v = std::ldexp(1.0f + m / 8.0f, e - 7);
    t[c] = (c & 0x80) ? -v : v;
  }
}

// X [n][d] fp32 -> codes [n][d] (allocated here) and power-of-two column
// scales [d]; uploads the decoding table.  0 or an error code (err set).
int fp8_quantize(const float* x, int64_t n, int d, uint8_t** q, float** scale, hipStream_t s,
                 std::string* err) {
  float table[256];
  fill_e4m3_table(table);
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(kE4M3), table, sizeof(table));
  if (e == hipSuccess && !*q) e = hipMalloc(q, (size_t)n * d);
  if (e == hipSuccess && !*scale) e = hipMalloc(scale, sizeof(float) * d);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(fp8_column_scales, dim3(d), dim3(256), 0, s, x, n, d, *scale);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    const int64_t total = n * (int64_t)d;
    const int blocks = (int)std::min<int64_t>(4096, (total + 255) / 256);
    hipLaunchKernelGGL(fp8_encode, dim3(blocks), dim3(256), 0, s, x, total, d, *scale, *q);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    *err = std::string("fp8 quantisation: ") + hipGetErrorString(e);
    return NAX_EHIP;
  }
  return NAX_OK;
}

Potential* make_logistic_mfma(const float* x, const float* y, int64_t n, int d,
                              double prior_scale, bool bf16, bool single, hipStream_t s,
                              std::string* err);

// fuse_variant: 0 = auto (matrix cores when num_chains >= 32), 1 = VALU tile
// kernel, 2 = matrix cores, 3 = a group of wavefronts per chain.  bf16 always runs on
// the matrix cores: 0 or 4 keep one X copy (bf16 residuals in X^T R), 2 keeps
// a second, fp16 copy for X^T R with fp16 residuals.
Potential* make_logistic(const float* x, const float* y, int64_t n, int d, double prior_scale,
                         int x_dtype, int fuse_variant, int num_chains, hipStream_t s,
                         std::string* err) {
  if (x_dtype == NAX_X_BF16) {
    if (fuse_variant != 0 && fuse_variant != 2 && fuse_variant != 4) {
      *err = "bf16 storage takes kernel 0 or 4 (one copy of X) or 2 (two copies)";
      return nullptr;
    }
    return make_logistic_mfma(x, y, n, d, prior_scale, true, fuse_variant != 2, s, err);
  }
  if (fuse_variant == 4) {
    *err = "kernel 4 (single-copy bf16) needs bf16 storage";
    return nullptr;
  }
  if (x_dtype == NAX_X_FP8 && fuse_variant != 0 && fuse_variant != 1) {
    *err = "fp8 storage runs on the VALU tile kernel: kernel 0 (auto) or 1";
    return nullptr;
  }
  if (x_dtype != NAX_X_F32 && x_dtype != NAX_X_FP8) {
    *err = "x_dtype not supported";
    return nullptr;
  }
  // a handful of chains on a small design: a group of wavefronts per chain,
  // the whole sampling loop in one persistent launch
  const bool fp8 = x_dtype == NAX_X_FP8;
  if (!fp8 && (fuse_variant == 3 || (fuse_variant == 0 && num_chains <= 16 && d <= 64 &&
                                     (double)n * d <= (double)(1 << 20))))
    return make_logistic_chain(x, y, n, d, prior_scale, s, err);
  const bool mfma = !fp8 && (fuse_variant == 2 ||
                             (fuse_variant == 0 && num_chains >= 32 && d <= 256));
  if (mfma && !(fuse_variant == 0 && num_chains <= 128 && d <= 320))
    return make_logistic_mfma(x, y, n, d, prior_scale, false, false, s, err);
  if (d > 320) {
    *err = "the fp32 logistic tile kernel supports dim <= 320";
    return nullptr;
  }
  auto* p = new LogisticPotential();
  p->x = x;
  p->y = y;
  p->n = n;
  p->d = d;
  p->prior_scale = prior_scale;
  p->x_dtype = x_dtype;
  if (fp8) {
    // the library's own e4m3 copy of X with power-of-two column scales
    p->x = nullptr;
    const int qrc = fp8_quantize(x, n, d, &p->q8, &p->scale8, s, err);
    if (qrc) {
      delete p;
      return nullptr;
    }
  }
  if (fuse_variant != 0 || num_chains < 1) return p;
  // Autotuner ("auto" kernel): row splits of the tile kernel, and for 32 <=
  // C <= 128 the tile kernel against the fp32 MFMA kernel, each timed over a
  // few evaluations of all C chains at z = 0; the fastest is kept.
  const int C = num_chains;
  const int ld = (d + 63) / 64 * 64;
  float *z = nullptr, *g = nullptr;
  double* pe = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool ok = hipMalloc(&z, sizeof(float) * (size_t)C * ld) == hipSuccess &&
            hipMalloc(&g, sizeof(float) * (size_t)C * ld) == hipSuccess &&
            hipMalloc(&pe, sizeof(double) * C) == hipSuccess &&
            hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
            hipMemsetAsync(z, 0, sizeof(float) * (size_t)C * ld, s) == hipSuccess;
  auto time_it = [&](Potential* q) -> float {
    if (q->eval(s, z, ld, C, pe, g, ld) != hipSuccess) return 1e30f;  // warm-up, workspaces
    hipEventRecord(e0, s);
    for (int r = 0; r < 3; ++r) q->eval(s, z, ld, C, pe, g, ld);
    hipEventRecord(e1, s);
    float ms = 1e30f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      return 1e30f;
    return ms;
  };
  Potential* best = p;
  if (ok) {
    const int64_t ntiles = (n + 31) / 32;
    const int ctiles = (C + 7) / 8;
    float best_ms = 1e30f;
    int best_rb = 0;
    for (int cand : {2048, 1024, 512, 256, 128, 64}) {
      const int rbv = (int)std::max<int64_t>(1, std::min<int64_t>(ntiles, cand / ctiles));
      p->rb_pick = rbv;
      const float ms = time_it(p);
      if (ms < best_ms) { best_ms = ms; best_rb = rbv; }
    }
    p->rb_pick = best_rb;
    if (mfma) {
      std::string e2;
      Potential* q = make_logistic_mfma(x, y, n, d, prior_scale, false, false, s, &e2);
      if (q) {
        if (time_it(q) < best_ms) best = q;
        else delete q;
      }
    }
  }
  hipFree(z); hipFree(g); hipFree(pe);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipStreamSynchronize(s);
  if (best != p) delete p;
  return best;
}

}  // namespace nax

extern "C" int nax_logistic_potential_grad_noalloc(const float* d_x, const float* d_y, int64_t n,
                                                   int32_t d, double prior_scale,
                                                   const float* d_z, int32_t ld, int32_t count,
                                                   double* d_pe, float* d_grad, void* stream) {
  using namespace nax;
  if (!d_x || !d_y || !d_z || !d_pe || !d_grad || n < 1 || d < 1 || count < 1 || ld < d) {
    set_error("nax_logistic_potential_grad_noalloc: need x, y, z, pe, grad, n >= 1, d >= 1, "
              "count >= 1, ld >= d");
    return NAX_EINVAL;
  }
  if (d > 320) {
    set_error("nax_logistic_potential_grad_noalloc: d <= 320 (the fp32 tile kernel)");
    return NAX_EINVAL;
  }
  if (!(prior_scale > 0.0) || !std::isfinite(prior_scale)) {
    set_error("nax_logistic_potential_grad_noalloc: prior_scale must be finite and > 0");
    return NAX_EINVAL;
  }
  const int RB = noalloc_row_splits(n, d, count);
  if (RB < 1) {
    set_error("nax_logistic_potential_grad_noalloc: count * (8 + 4 d) bytes of partials exceed "
              "the 16 MiB scratch array (count * d at most about 4.19e6)");
    return NAX_EINVAL;
  }
  unsigned char* base = nullptr;
  hipError_t e = scratch_address(&base);
  if (e != hipSuccess) {
    set_error(std::string("nax_logistic_potential_grad_noalloc: hipGetSymbolAddress: ") +
              hipGetErrorString(e));
    return NAX_EHIP;
  }
  double* wsu = reinterpret_cast<double*>(base);
  const size_t ubytes = ((size_t)RB * count * sizeof(double) + 255) / 256 * 256;
  float* wsg = reinterpret_cast<float*>(base + ubytes);
  hipStream_t s = static_cast<hipStream_t>(stream);
  constexpr int BN = 32, BC = 8;
  const int ctiles = (count + BC - 1) / BC;
  const int dch = (d + 63) / 64;
  const int DP = dch * 64 + 1;
  const size_t shm =
      sizeof(float) * ((size_t)BC * DP + (size_t)BN * DP + BN * BC) + 256 * sizeof(double);
  dim3 grid(RB, ctiles);
#define NAX_LAUNCH(DC)                                                                      \
  hipLaunchKernelGGL((logit_partial_f32<BN, BC, DC, XF32>), grid, dim3(256), shm, s, XF32{d_x},  \
                     d_y, n, d, d_z, ld, count, RB, wsg, wsu)
  switch (dch) {
    case 1: NAX_LAUNCH(1); break;
    case 2: NAX_LAUNCH(2); break;
    case 3: NAX_LAUNCH(3); break;
    case 4: NAX_LAUNCH(4); break;
    case 5: NAX_LAUNCH(5); break;
  }
#undef NAX_LAUNCH
  e = hipGetLastError();
  if (e == hipSuccess) {
    const double inv_s2 = 1.0 / (prior_scale * prior_scale);
    const double log_norm = std::log(std::sqrt(2.0 * M_PI) * prior_scale);
    hipLaunchKernelGGL(logit_reduce, dim3(count), dim3(256), 0, s, wsg, wsu, RB, count, d, d_z,
                       ld, inv_s2, log_norm, 0.0, d_pe, d_grad, ld, nullptr, nullptr, 0, 0);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    set_error(std::string("nax_logistic_potential_grad_noalloc: ") + hipGetErrorString(e));
    return NAX_EHIP;
  }
  return NAX_OK;
}

extern "C" int nax_fp8_quantize(const float* d_x, int64_t n, int32_t d, uint8_t* d_codes,
                                float* d_scales) {
  using namespace nax;
  if (!d_x || !d_codes || !d_scales || n < 1 || d < 1) {
    set_error("nax_fp8_quantize: need x, codes, scales, n >= 1, d >= 1");
    return NAX_EINVAL;
  }
  std::string err;
  const int rc = fp8_quantize(d_x, n, d, &d_codes, &d_scales, nullptr, &err);
  if (rc) set_error(err);
  return rc;
}
