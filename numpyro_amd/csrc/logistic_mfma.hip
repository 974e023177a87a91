// Logistic potential + adjoint on the matrix cores (gfx950 MFMA), C chains.
//
// Same mathematics as logistic.hip (numpyro/distributions/discrete.py:244-263,
// distributions/util.py:317-320, examples/covtype.py:66-71):
//   L = X Z^T,  R = sigmoid(L) - y,  U_c = sum_n softplus(L) - y L,  G^T = R^T X.
//
// Work split: a block is 4 waves; a wave owns 32 chains and walks 32-row tiles
// of X.  Per tile:
//   GEMM1  L[32 rows x 32 chains] = sum_k X[., k] Z[., k]   (A = X, B = Z^T)
//   epilogue on the 16 accumulator values each lane holds
//   GEMM2  G^T[32 chains x D] += R^T[32 x 32 rows] X[32 rows x D]
// The accumulator of GEMM1 holds, in lane (c, h) and register v, L at row
// 8(v/4) + 4h + v%4 of chain c.  GEMM2 takes R^T as its A operand, whose lane
// (c, h) must supply K-slots 8h..8h+7 of chain c: choosing the K order of
// GEMM2 so that slot 8h + j of MFMA m is row 16m + 8(j/4) + 4h + j%4 makes that
// operand exactly accumulator registers 8m..8m+7 -- R never leaves registers.
// X is stored twice in fragment order (prepared once per model):
//   A1[T][kk][lane][j] = X[32T + lane%32][KS kk + KS/2 (lane/32) + j]
//   B2[T][m][dt][lane][j] = X[32T + row(m, lane/32, j)][32 dt + lane%32]
// so every operand fetch is one coalesced load.  Blocks that walk the same
// row tiles (same row split, different chain tiles) are placed on the same
// XCD (blocks are dealt to the 8 XCDs round-robin) to share its L2.
//
// Precision forms:
//   BF16: X in bf16 (RNE); Z enters as a bf16 hi+lo pair (two MFMAs), so the
//         positions keep ~16 mantissa bits; R as bf16 hi+lo; fp32 accumulate.
//   FP32: X, Z, R in fp32 on the fp32 MFMA (v_mfma_f32_32x32x2_f32).
// Energies accumulate in fp64 per lane.  Partials per row split are summed in
// a fixed order by logit_reduce (logistic.hip).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>

#include "nax_device.h"
#include "nax_internal.h"

namespace nax {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__device__ __forceinline__ float softplus_f(float x) {
  return fmaxf(x, 0.0f) + log1pf(__expf(-fabsf(x)));
}
__device__ __forceinline__ float sigmoid_f(float x) { return 1.0f / (1.0f + __expf(-x)); }

// r = sigmoid(l) - y and u = softplus(l) - y l (logistic_terms_f) with the
// sign test shared: r = (l >= 0 ? 1 : e) / (1 + e) - y as one fma, and
// max(l, 0) as a select on the same test.
__device__ __forceinline__ void logit_terms(float l, float y, float& r, float& u) {
  const bool pos = l >= 0.0f;
  const float e = __builtin_amdgcn_exp2f(-fabsf(l) * 1.44269504088896341f);
  const float d = 1.0f + e;
  const float inv = __builtin_amdgcn_rcpf(d);
  r = fmaf(pos ? 1.0f : e, inv, -y);
  u = fmaf(__builtin_amdgcn_logf(d), 0.693147180559945309f, fmaf(-y, l, pos ? l : 0.0f));
}


// ---- fragment-order preparation -------------------------------------------------
// KS = K per MFMA (16 for bf16, 2 for fp32); Dp = D rounded up to 32.
// SWZ (single-copy bf16 form): within each 64-fragment block the second
// half's rows are stored XOR 4 (position p holds lane p ^ 4 for p >= 32), so
// the transposed GEMM2 reads of a row's two 8-column halves fall on
// different LDS banks.
template <typename T, int KS, bool SWZ = false>
__global__ void prep_a1(const float* __restrict__ X, int64_t N, int D, int Dp, int64_t ntile,
                        T* __restrict__ out) {
  const int KK = Dp / KS;
  const int64_t total = ntile * KK * 64 * (KS / 2);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int j = (int)(i % (KS / 2));
    int64_t r = i / (KS / 2);
    int lane = (int)(r % 64);
    if (SWZ) lane ^= (lane >> 5) << 2;
    r /= 64;
    int kk = (int)(r % KK);
    int64_t t = r / KK;
    int64_t n = 32 * t + lane % 32;
    int d = KS * kk + (KS / 2) * (lane / 32) + j;
    float v = (n < N && d < D) ? X[n * D + d] : 0.0f;
    out[i] = (T)v;
  }
}

template <typename T, int KS, bool VIA_BF16 = false>
__global__ void prep_b2(const float* __restrict__ X, int64_t N, int D, int Dp, int64_t ntile,
                        T* __restrict__ out) {
  const int DT = Dp / 32;
  const int M = 32 / KS;  // MFMAs per 32-row tile in GEMM2
  const int64_t total = ntile * M * DT * 64 * (KS / 2);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int j = (int)(i % (KS / 2));
    int64_t r = i / (KS / 2);
    int lane = (int)(r % 64);
    r /= 64;
    int dt = (int)(r % DT);
    r /= DT;
    int m = (int)(r % M);
    int64_t t = r / M;
    int h = lane / 32;
    // K-slot (KS/2) h + j of MFMA m <-> accumulator register v = (KS/2) m + j
    int v = (KS / 2) * m + j;
    int row = 8 * (v / 4) + 4 * h + (v % 4);
    int64_t n = 32 * t + row;
    int d = 32 * dt + lane % 32;
    float val = (n < N && d < D) ? X[n * D + d] : 0.0f;
    if (VIA_BF16) val = (float)(__bf16)val;  // the same quantised X as the A1 copy
    out[i] = (T)val;
  }
}

__global__ void pad_y(const float* __restrict__ y, int64_t N, int64_t Np, float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < Np;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = i < N ? y[i] : 0.5f;
}

// ---- the fused kernel ----------------------------------------------------------------
struct FormBF16 {
  static constexpr int KS = 16;
  typedef bf16x8 frag;
  static __device__ __forceinline__ f32x16 mma(frag a, frag b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};
struct FormF32 {
  static constexpr int KS = 2;
  typedef float frag;
  static __device__ __forceinline__ f32x16 mma(frag a, frag b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
};

template <typename F, int DP>
__global__ void __launch_bounds__(256, 1)
logit_mfma(const typename F::frag* __restrict__ A1, const typename F::frag* __restrict__ B2,
           const float* __restrict__ ypad, int64_t N, int64_t ntile, const float* __restrict__ Z,
           int ldz, int D, int C, float* __restrict__ wsg,
           double* __restrict__ wsu, const int* __restrict__ act, const int* __restrict__ nact) {
  constexpr int KS = F::KS;
  constexpr int KK = DP / KS;   // GEMM1 K steps
  constexpr int DT = DP / 32;   // GEMM2 d tiles
  constexpr int M2 = 32 / KS;   // GEMM2 MFMAs per d tile and row tile
  constexpr bool BF = (KS == 16);
  typedef typename F::frag frag;
  // one row tile in LDS: A1 part, B2 part, 32 labels
  constexpr int ABYTES = KK * 64 * (int)sizeof(frag);
  constexpr int BBYTES = M2 * DT * 64 * (int)sizeof(frag);
  constexpr int TBYTES = ABYTES + BBYTES + 128;
  constexpr int PF = (ABYTES + BBYTES) / (256 * 16);   // 16-byte chunks per thread
  constexpr int PA = ABYTES / 16;                      // chunks in the A1 part
  static_assert((ABYTES + BBYTES) % (256 * 16) == 0, "tile must split into 16-byte chunks");
  __shared__ __attribute__((aligned(16))) char lds[2][TBYTES];

  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int l = tid & 63;
  const int cl = l & 31, h = l >> 5;
  // XCD-aware block placement: blocks of one row split share an XCD.
  // Chains are taken through the compacted list act[0..*nact) of the chains
  // that requested a gradient (all C in order when act is null); column
  // index ci addresses the workspaces (row stride CT * 128), chain act[ci]
  // the positions.  The chain tiles and row splits are laid out for the
  // requesting chains only, over the whole launch.
  const int n_act = act ? *nact : C;
  if (n_act <= 0) return;
  int CT, RB;
  logit_split(n_act, (int)gridDim.x, ntile, CT, RB);
  const int ldw = CT * 128;
  const int bid = blockIdx.x;
  const int xcd = bid & 7;
  const int slot = bid >> 3;
  const int ct = slot % CT;
  const int rb = (slot / CT) * 8 + xcd;
  if (rb >= RB) return;  // uniform over the block
  const int c0 = ct * 128 + wave * 32;
  const bool cvalid = c0 + cl < n_act;
  const int c = cvalid ? (act ? act[c0 + cl] : c0 + cl) : 0;

  // Z fragments (B operand of GEMM1): lane (c, h) holds Z[c][KS kk + KS/2 h + j]
  frag zh[KK];
  frag zlo[BF ? KK : 1];
#pragma unroll
  for (int kk = 0; kk < KK; ++kk) {
    if constexpr (BF) {
      bf16x8 hi, lo;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        int d = 16 * kk + 8 * h + j;
        float v = (cvalid && d < D) ? Z[(size_t)c * ldz + d] : 0.0f;
        __bf16 b = (__bf16)v;
        hi[j] = b;
        lo[j] = (__bf16)(v - (float)b);
      }
      zh[kk] = hi;
      zlo[kk] = lo;
    } else {
      int d = 2 * kk + h;
      zh[kk] = (cvalid && d < D) ? Z[(size_t)c * ldz + d] : 0.0f;
    }
  }

  f32x16 g[DT];
#pragma unroll
  for (int dt = 0; dt < DT; ++dt)
#pragma unroll
    for (int v = 0; v < 16; ++v) g[dt][v] = 0.0f;
  CompSum u;  // fp64 energy over this lane's tiles, compensated

  // ---- tile staging: global -> registers (prefetch) -> LDS ----
  // Clang vector types, not HIP's uint4/float4 structs: an array of those
  // stays in scratch memory, and every prefetch then waits for its loads.
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  u32x4 pf[PF];
  f32x4 pfy;
  auto load_tile = [&](int64_t t) {
    const u32x4* ga = (const u32x4*)(A1 + (size_t)t * KK * 64);
    const u32x4* gb = (const u32x4*)(B2 + (size_t)t * M2 * DT * 64);
#pragma unroll
    for (int p = 0; p < PF; ++p) {
      const int idx = tid + 256 * p;
      if constexpr (PA % 256 == 0) {  // the source of chunk p is known at compile time
        if (p < PA / 256) pf[p] = ga[idx];
        else pf[p] = gb[idx - PA];
      } else {
        pf[p] = idx < PA ? ga[idx] : gb[idx - PA];
      }
    }
    pfy = ((const f32x4*)(ypad + 32 * t))[tid & 7];
  };
  auto store_tile = [&](int buf) {
    u32x4* s4 = (u32x4*)lds[buf];
#pragma unroll
    for (int p = 0; p < PF; ++p) s4[tid + 256 * p] = pf[p];
    // every thread stores its copy of label chunk tid & 7 (equal values):
    // an unconditional store keeps the label load at the top of the tile
    ((f32x4*)(lds[buf] + ABYTES + BBYTES))[tid & 7] = pfy;
  };

  int64_t t = rb;
  int cur = 0;
  if (t < ntile) {
    load_tile(t);
    store_tile(0);
  }
  __syncthreads();
  for (; t < ntile; t += RB) {
    const bool has_next = t + RB < ntile;
    // unconditional (the last tile reloads itself, unused), so that the
    // prefetch registers are written on every path and stay registers
    load_tile(has_next ? t + RB : t);
    // keep the prefetch loads here: left to itself the scheduler sinks them
    // next to their LDS stores at the end of the tile and waits on them
    __builtin_amdgcn_sched_barrier(0);
    const frag* la = (const frag*)lds[cur];
    const frag* lb = (const frag*)(lds[cur] + ABYTES);
    const float* ly = (const float*)(lds[cur] + ABYTES + BBYTES);

    // GEMM1: L = X_tile Z^T
    f32x16 L;
#pragma unroll
    for (int v = 0; v < 16; ++v) L[v] = 0.0f;
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
      frag af = la[kk * 64 + l];
      L = F::mma(af, zh[kk], L);
      if constexpr (BF) L = F::mma(af, zlo[kk], L);
    }
    // epilogue: register v holds row 8(v/4) + 4h + v%4 (padded rows give
    // r = 0 and u = fp32(ln 2), removed by the reduction)
    float rv[16];
    float us = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 yq = *(const float4*)(ly + 8 * q + 4 * h);
      const float yy[4] = {yq.x, yq.y, yq.z, yq.w};
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        float uu;
        logit_terms(L[4 * q + m], yy[m], rv[4 * q + m], uu);
        us += uu;
      }
    }
    u.add((double)us);
    // GEMM2: G^T += R^T X_tile
    if constexpr (BF) {
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        bf16x8 rh, rl;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          __bf16 hb = (__bf16)rv[8 * m + j];
          rh[j] = hb;
          rl[j] = (__bf16)(rv[8 * m + j] - (float)hb);
        }
#pragma unroll
        for (int dt = 0; dt < DT; ++dt) {
          frag bf = lb[(m * DT + dt) * 64 + l];
          g[dt] = F::mma(rh, bf, g[dt]);
          g[dt] = F::mma(rl, bf, g[dt]);
        }
      }
    } else {
#pragma unroll
      for (int m = 0; m < 16; ++m) {
#pragma unroll
        for (int dt = 0; dt < DT; ++dt) {
          frag bf = lb[(m * DT + dt) * 64 + l];
          g[dt] = F::mma(rv[m], bf, g[dt]);
        }
      }
    }
    store_tile(cur ^ 1);  // unconditional: the other buffer is not read after the last tile
    __syncthreads();
    cur ^= 1;
  }

  // G^T partial: register v of tile dt is chain 8(v/4) + 4h + v%4 (relative
  // to c0), column d = 32 dt + cl.
#pragma unroll
  for (int dt = 0; dt < DT; ++dt) {
    const int d = 32 * dt + cl;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int ci = c0 + 8 * (v / 4) + 4 * h + (v % 4);
      if (ci < n_act && d < D) wsg[((size_t)rb * ldw + ci) * D + d] = g[dt][v];
    }
  }
  CompSum uo;
  uo.s = __shfl_xor(u.s, 32, 64);
  uo.c = __shfl_xor(u.c, 32, 64);
  u.add(uo);
  const double ut = u.value();
  if (h == 0 && cvalid) wsu[(size_t)rb * ldw + c0 + cl] = ut;
}

// ---- pipelined BF16 kernel --------------------------------------------------------
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;

// One 16-byte global->LDS load per lane (global_load_lds_dwordx4): lane l
// writes LDS byte lds_base + 16 l (lds_base wave-uniform, in M0).  Issued as
// inline assembly so that the compiler does not tie later LDS reads to it:
// given the builtin, the waitcnt pass makes every later ds_read of the block
// wait for all outstanding LDS-DMA loads (vmcnt(0)), which turns a prefetch
// several tiles ahead into a wait on the tile just issued.  The kernel waits
// for these loads itself with counted s_waitcnt vmcnt + a barrier.
__device__ __forceinline__ void lds_dma16(const void* gptr, uint32_t lds_base) {
  asm volatile(
      "s_mov_b32 m0, %1\n\t"
      "s_nop 0\n\t"
      "global_load_lds_dwordx4 %0, off"
      :
      : "v"(gptr), "s"(lds_base)
      : "memory", "m0");
}
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

// Software-pipelined form of logit_mfma for bf16 X.  The block's row tiles
// i = 0..n-1 (tile rb + i RB) stream through LDS by direct global->LDS loads
// (no staging registers).  Iteration i runs, as one branch-free basic block,
// the epilogue of tile i (VALU) interleaved in source order with GEMM1 of
// tile i+1 and GEMM2 of tile i-1 (MFMA), so the matrix cores and the VALU
// work on different tiles at once.  The three parts of a tile live for
// different iterations (A1 of tile t is read in iteration t-1, y in t, B2 in
// t+1), so each has its own 4-slot ring: iteration i issues A1 of tile i+4,
// B2 of tile i+2 and y of tile i+3 into the slots freed by iteration i-1,
// and then waits only for the loads issued two iterations earlier -- every
// load has three iterations (~4 us) to arrive.
//
// SINGLE = false (two-copy form): GEMM2 runs in fp16 on the B2 copy: X (bf16
// values) is exact in fp16 and R gets 11 mantissa bits in a single MFMA.
// SINGLE = true (single-copy form, the default): there is no B2; GEMM2 runs
// in bf16 and reads its X operand transposed from the A1 tile, which
// therefore stays in LDS until iteration t+1 (8-slot A1 ring), and R gets
// 8 mantissa bits.
template <int DP, bool SINGLE>
__global__ void __launch_bounds__(256, 1)
logit_bf16_pipe(const bf16x8* __restrict__ A1, const f16x8* __restrict__ B2,
                const float* __restrict__ ypad, int64_t N, int64_t ntile,
                const float* __restrict__ Z, int ldz, int D, int C,
                float* __restrict__ wsg, double* __restrict__ wsu,
                const int* __restrict__ act, const int* __restrict__ nact) {
  constexpr int KK = DP / 16;
  constexpr int DT = DP / 32;
  constexpr int ABYTES = KK * 64 * 16;
  constexpr int BBYTES = 2 * DT * 64 * 16;
  constexpr int TBYTES = ABYTES + BBYTES + 128;
  // A1 and B2 parts of a tile are KK wave-loads of 64 16-byte chunks each;
  // wave w issues the wave-loads w, w + 4, ... of each part.
  constexpr int PP = (KK + 3) / 4;
  static_assert(ABYTES == KK * 64 * 16 && BBYTES == KK * 64 * 16, "part sizes");
  // loads per iteration of the wave that issues the fewest (wave 3)
  constexpr int NMIN = (SINGLE ? 1 : 2) * (KK / 4) + 1;
  (void)TBYTES;
  // SINGLE: GEMM2 reads X transposed from the A1 tile, which then lives
  // from iteration t-1 (GEMM1) to t+1 (GEMM2): 8 A1 slots, no B2 ring.
  constexpr int NA = SINGLE ? 8 : 4;
  __shared__ __attribute__((aligned(16))) char lds_a[NA][ABYTES];
  __shared__ __attribute__((aligned(16))) char lds_b[SINGLE ? 1 : 4][SINGLE ? 16 : BBYTES];
  __shared__ __attribute__((aligned(16))) char lds_y[4][128];
  typedef typename std::conditional<SINGLE, bf16x8, f16x8>::type rfrag;

  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int l = tid & 63;
  const int cl = l & 31, h = l >> 5;
  // Chains are taken through the compacted list act[0..*nact) of the chains
  // that requested a gradient (all C in order when act is null); column
  // index ci addresses the workspaces (row stride CT * 128), chain act[ci]
  // the positions.  The chain tiles and row splits are laid out for the
  // requesting chains only, over the whole launch.
  const int n_act = act ? *nact : C;
  if (n_act <= 0) return;
  int CT, RB;
  logit_split(n_act, (int)gridDim.x, ntile, CT, RB);
  const int ldw = CT * 128;
  const int bid = blockIdx.x;
  const int xcd = bid & 7;
  const int slot = bid >> 3;
  const int ct = slot % CT;
  const int rb = (slot / CT) * 8 + xcd;
  if (rb >= RB) return;  // uniform over the block
  const int c0 = ct * 128 + wave * 32;
  const bool cvalid = c0 + cl < n_act;
  const int c = cvalid ? (act ? act[c0 + cl] : c0 + cl) : 0;

  bf16x8 zh[KK], zlo[KK];
#pragma unroll
  for (int kk = 0; kk < KK; ++kk) {
    bf16x8 hi, lo;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int d = 16 * kk + 8 * h + j;
      const float v = (cvalid && d < D) ? Z[(size_t)c * ldz + d] : 0.0f;
      const __bf16 b = (__bf16)v;
      hi[j] = b;
      lo[j] = (__bf16)(v - (float)b);
    }
    zh[kk] = hi;
    zlo[kk] = lo;
  }
  f32x16 g[DT];
#pragma unroll
  for (int dt = 0; dt < DT; ++dt)
#pragma unroll
    for (int v = 0; v < 16; ++v) g[dt][v] = 0.0f;
  CompSum u;  // fp64 energy over this lane's tiles, compensated

  const int64_t n = rb < ntile ? (ntile - 1 - rb) / RB + 1 : 0;
  // Tile i is clamped to the block's last tile, so every load is in bounds
  // and a wave issues the same number of loads in every iteration, which
  // the counted waits below rely on.  Lane l of wave w loads chunk
  // tid + 256 p to LDS offset 16 (tid + 256 p): each wave's 64 chunks are
  // contiguous, as the LDS-direct load (LDS address = wave base + 16 lane)
  // requires.  The 32 labels of a tile are 8 chunks: lanes 0-1 of each wave.
  auto tile_of = [&](int64_t i) { return rb + (i < n ? i : n - 1) * (int64_t)RB; };
  auto issue_a = [&](int64_t i) {
    const uint4* ga = (const uint4*)(A1 + (size_t)tile_of(i) * KK * 64);
    char* base = lds_a[i & (NA - 1)];
#pragma unroll
    for (int p = 0; p < PP; ++p)
      if (KK % 4 == 0 || 4 * p + wave < KK)
        lds_dma16(ga + tid + 256 * p, lds_addr(base + 16 * (256 * p + 64 * wave)));
  };
  auto issue_b = [&](int64_t i) {
    if constexpr (!SINGLE) {
      const uint4* gb = (const uint4*)(B2 + (size_t)tile_of(i) * 2 * DT * 64);
      char* base = lds_b[i & 3];
#pragma unroll
      for (int p = 0; p < PP; ++p)
        if (KK % 4 == 0 || 4 * p + wave < KK)
          lds_dma16(gb + tid + 256 * p, lds_addr(base + 16 * (256 * p + 64 * wave)));
    }
  };
  auto issue_y = [&](int64_t i) {
    if (l < 2) lds_dma16(ypad + 32 * tile_of(i) + 4 * (2 * wave + l), lds_addr(lds_y[i & 3] + 32 * wave));
  };
  auto issue_group = [&](int64_t i) {
    issue_a(i + 4);
    issue_b(i + 2);
    issue_y(i + 3);
  };
  // After iteration i only the groups of iterations i-1 and i may still be
  // in flight (at most 2 NMIN loads of any wave, loads retire in order); the
  // barrier then publishes every wave's loads to the block.
  static_assert(2 * NMIN < 64, "vmcnt holds 6 bits");
  auto wait_group = [&]() {
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(2 * NMIN) : "memory");
    __syncthreads();
  };
  auto wait_all = [&]() {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  };
  // GEMM2's B operand for MFMA (m2, dt) of tile i: X rows 16 m2 + 8 (j / 4)
  // + 4 h + j % 4 (the K order that makes R^T the GEMM1 accumulator
  // registers), column 32 dt + lane % 32.  Two-copy form: one 16-byte read
  // of the fragment-ordered B2 copy.  Single-copy form: two
  // ds_read_b64_tr_b16 of the A1 tile, whose rows hold 8 consecutive
  // columns.  In each 16-lane group, lane q of the transposed read returns
  // element q % 4 of the 8-byte chunks addressed by lanes q / 4 + 4 k
  // (k = 0..3), so lane q addresses row (q / 4) of its group's 4-row block at
  // columns 4 (q % 4) .. 4 (q % 4) + 3 of its group's 16-column half.
  // The single-copy A1 stores the second half's rows XOR 4 (prep_a1<SWZ>):
  // GEMM1 lane l reads its fragment at position lsw, and the transposed
  // reads address row r of half h at position 32 h + (r ^ 4 h).
  const int lsw = SINGLE ? (l ^ ((l >> 5) << 2)) : l;
  const int jl = l & 15, gl = l >> 4;
  const int hb = (jl & 3) >> 1;
  const int tr_off =
      (((gl & 1) * 64 + hb * 32 + 4 * ((gl >> 1) ^ (SINGLE ? hb : 0)) + (jl >> 2)) * 16) +
      8 * (jl & 1);
  auto bfrag = [&](int64_t i, int m2, int dt) -> rfrag {
    if constexpr (SINGLE) {
      typedef short v4s __attribute__((ext_vector_type(4)));
      typedef short v8s __attribute__((ext_vector_type(8)));
      const char* base = lds_a[i & (NA - 1)] + tr_off + dt * 2048 + m2 * 256;
      const v4s lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
          (__attribute__((address_space(3))) v4s*)(base));
      const v4s hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
          (__attribute__((address_space(3))) v4s*)(base + 128));
      const v8s w = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
      return __builtin_bit_cast(rfrag, w);
    } else {
      return ((const f16x8*)lds_b[i & 3])[(m2 * DT + dt) * 64 + l];
    }
  };
  auto mma2 = [&](rfrag a, rfrag b, const f32x16& acc) -> f32x16 {
    if constexpr (SINGLE) return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc, 0, 0, 0);
  };
  auto gemm1 = [&](int64_t i) {
    const bf16x8* la = (const bf16x8*)lds_a[i & (NA - 1)];
    f32x16 L;
#pragma unroll
    for (int v = 0; v < 16; ++v) L[v] = 0.0f;
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
      const bf16x8 af = la[kk * 64 + lsw];
      L = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af, zh[kk], L, 0, 0, 0);
      L = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af, zlo[kk], L, 0, 0, 0);
    }
    return L;
  };
  auto gemm2 = [&](int64_t i, rfrag r0, rfrag r1) {
#pragma unroll
    for (int dt = 0; dt < DT; ++dt) {
      g[dt] = mma2(r0, bfrag(i, 0, dt), g[dt]);
      g[dt] = mma2(r1, bfrag(i, 1, dt), g[dt]);
    }
  };
  // Rows past N have X = 0 and y = 1/2, so r = 0 exactly and u = fp32(ln 2)
  // (removed by the reduction); chains past C have Z = 0 and are never
  // written.  No per-element masking is needed.  Register v of L holds row
  // 8(v/4) + 4h + v%4.
  auto load_y = [&](int64_t i, float* yv) {
    const float* ly = (const float*)lds_y[i & 3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 yq = *(const float4*)(ly + 8 * q + 4 * h);
      yv[4 * q] = yq.x;
      yv[4 * q + 1] = yq.y;
      yv[4 * q + 2] = yq.z;
      yv[4 * q + 3] = yq.w;
    }
  };
  auto epilogue = [&](int64_t i, const f32x16& L, rfrag& r0, rfrag& r1) {
    float yv[16];
    load_y(i, yv);
    float us = 0.0f;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      float r, uu;
      logit_terms(L[v], yv[v], r, uu);
      us += uu;
      if (v < 8) r0[v] = (typename std::remove_reference<decltype(r0[0])>::type)r;
      else r1[v - 8] = (typename std::remove_reference<decltype(r1[0])>::type)r;
    }
    u.add((double)us);
  };
  // steady-state step: epilogue(i) || GEMM1(i+1) || GEMM2(i-1), interleaved
  // element by element in source order.
  auto fused_step = [&](int64_t i, const f32x16& Lc, rfrag rp0, rfrag rp1, rfrag& r0,
                        rfrag& r1) {
    const bf16x8* la = (const bf16x8*)lds_a[(i + 1) & (NA - 1)];
    float yv[16];
    load_y(i, yv);
    f32x16 Ln;
#pragma unroll
    for (int v = 0; v < 16; ++v) Ln[v] = 0.0f;
    float us = 0.0f;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
#pragma unroll
      for (int kk = e * KK / 16; kk < (e + 1) * KK / 16; ++kk) {
        const bf16x8 af = la[kk * 64 + lsw];
        Ln = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af, zh[kk], Ln, 0, 0, 0);
        Ln = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af, zlo[kk], Ln, 0, 0, 0);
      }
#pragma unroll
      for (int j = e * 2 * DT / 16; j < (e + 1) * 2 * DT / 16; ++j) {
        const int m2 = j / DT, dt = j % DT;
        g[dt] = mma2(m2 ? rp1 : rp0, bfrag(i - 1, m2, dt), g[dt]);
      }
      float r, uu;
      logit_terms(Lc[e], yv[e], r, uu);
      us += uu;
      if (e < 8) r0[e] = (typename std::remove_reference<decltype(r0[0])>::type)r;
      else r1[e - 8] = (typename std::remove_reference<decltype(r1[0])>::type)r;
    }
    u.add((double)us);
    return Ln;
  };
  auto fused_step256 = [&](int64_t i, const f32x16& Lc, rfrag rp0, rfrag rp1, rfrag& r0,
                        rfrag& r1) {
    const bf16x8* la = (const bf16x8*)lds_a[(i + 1) & (NA - 1)];
    float yv[16];
    load_y(i, yv);
    f32x16 Ln;
#pragma unroll
    for (int v = 0; v < 16; ++v) Ln[v] = 0.0f;
    float us = 0.0f;
    // chunk e: 2 GEMM1 MFMAs (k-step e) + 1 GEMM2 MFMA (j = e) + the
    // epilogue of element e; fragments for chunk e + 1 are read from LDS one
    // chunk ahead.  Scheduling barriers keep each chunk's VALU work next to
    // its MFMAs so that it runs in their shadow.
    bf16x8 af = la[lsw];
    rfrag bf = bfrag(i - 1, 0, 0);
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      bf16x8 af_n;
      rfrag bf_n;
      if (e + 1 < 16) {
        af_n = la[(e + 1) * 64 + lsw];
        const int jn = e + 1, m2n = jn / DT, dtn = jn % DT;
        bf_n = bfrag(i - 1, m2n, dtn);
      }
      Ln = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af, zh[e], Ln, 0, 0, 0);
      {
        const int m2 = e / DT, dt = e % DT;
        g[dt] = mma2(m2 ? rp1 : rp0, bf, g[dt]);
      }
      Ln = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af, zlo[e], Ln, 0, 0, 0);
      float r, uu;
      logit_terms(Lc[e], yv[e], r, uu);
      us += uu;
      asm volatile("" : "+v"(us));  // keeps this element's loss term in the chunk
      if (e < 8) r0[e] = (typename std::remove_reference<decltype(r0[0])>::type)r;
      else r1[e - 8] = (typename std::remove_reference<decltype(r1[0])>::type)r;
      if (e + 1 < 16) {
        af = af_n;
        bf = bf_n;
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    u.add((double)us);
    return Ln;
  };

  if (n > 0) {
    // the groups of the virtual iterations -4..-1: A1 of tiles 0-3, B2 of
    // tiles 0-1, y of tiles 0-2
    issue_a(0);
    issue_a(1);
    issue_y(0);
    issue_a(2);
    issue_b(0);
    issue_y(1);
    issue_a(3);
    issue_b(1);
    issue_y(2);
    wait_all();
    f32x16 Lc = gemm1(0);
    rfrag rp0, rp1;
    {  // i = 0: epilogue(0) + gemm1(1)
      __syncthreads();  // A1 slot 0 is refilled next: every wave has read tile 0
      issue_group(0);
      rfrag r0, r1;
      epilogue(0, Lc, r0, r1);
      if (n > 1) Lc = gemm1(1);
      rp0 = r0;
      rp1 = r1;
      wait_group();
    }
    for (int64_t i = 1; i + 1 < n; ++i) {  // steady state, branch-free body
      issue_group(i);
      rfrag r0, r1;
      if constexpr (DP == 256) Lc = fused_step256(i, Lc, rp0, rp1, r0, r1);
      else Lc = fused_step(i, Lc, rp0, rp1, r0, r1);
      rp0 = r0;
      rp1 = r1;
      wait_group();
    }
    if (n > 1) {  // i = n-1: epilogue(n-1) + gemm2(n-2)
      rfrag r0, r1;
      epilogue(n - 1, Lc, r0, r1);
      gemm2(n - 2, rp0, rp1);
      rp0 = r0;
      rp1 = r1;
    }
    wait_all();  // B2 of tile n-1 (and the clamped tail loads) have landed
    gemm2(n - 1, rp0, rp1);
  }

#pragma unroll
  for (int dt = 0; dt < DT; ++dt) {
    const int d = 32 * dt + cl;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int ci = c0 + 8 * (v / 4) + 4 * h + (v % 4);
      if (ci < n_act && d < D) wsg[((size_t)rb * ldw + ci) * D + d] = g[dt][v];
    }
  }
  CompSum uo;
  uo.s = __shfl_xor(u.s, 32, 64);
  uo.c = __shfl_xor(u.c, 32, 64);
  u.add(uo);
  const double ut = u.value();
  if (h == 0 && cvalid) wsu[(size_t)rb * ldw + c0 + cl] = ut;
}

__global__ void abs_max_kernel(const float* __restrict__ X, int64_t n, unsigned* __restrict__ out) {
  float m = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(X[i]));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

// Fixed-order sum of the partials (+ Normal prior), defined in logistic.hip.
__global__ void logit_reduce(const float* __restrict__ wsg, const double* __restrict__ wsu, int RB,
                             int C, int D, const float* __restrict__ Z, int ldz, double inv_s2,
                             double log_norm, double u_offset, double* __restrict__ pe,
                             float* __restrict__ g, int ldg, const int* __restrict__ act,
                             const int* __restrict__ nact, int nblk, int64_t ntile);

struct LogisticMfma : Potential {
  int64_t n = 0, ntile = 0;
  int d = 0, dp = 0;
  double prior_scale = 1.0;
  bool bf16 = true;
  bool single = false;  // bf16 only: one X copy, GEMM2 transposed from it, bf16 R
  void* a1 = nullptr;
  void* b2 = nullptr;
  float* ypad = nullptr;
  float* wsg = nullptr;
  double* wsu = nullptr;
  size_t ws_elems = 0;
  int num_cu = 256;

  ~LogisticMfma() override {
    hipFree(a1); hipFree(b2); hipFree(ypad); hipFree(wsg); hipFree(wsu);
  }
  int dim() const override { return d; }
  const char* name() const override {
    return bf16 ? (single ? "logit_mfma_bf16_single" : "logit_mfma_bf16") : "logit_mfma_f32";
  }
  double alg_flops(int C) const override { return 4.0 * (double)n * d * C; }
  double alg_bytes(int C) const override { return (bf16 ? 2.0 : 4.0) * (double)n * d + 4.0 * n; }

  hipError_t setup(const float* x, const float* y, hipStream_t s) {
    ntile = (n + 31) / 32;
    dp = (d + 31) / 32 * 32;
    const size_t elt = bf16 ? 2 : 4;
    const size_t per = (size_t)ntile * 32 * dp * elt;
    hipError_t e = hipMalloc(&a1, per);
    if (e != hipSuccess) return e;
    if (!single) {
      e = hipMalloc(&b2, per);
      if (e != hipSuccess) return e;
    }
    e = hipMalloc(&ypad, sizeof(float) * ntile * 32);
    if (e != hipSuccess) return e;
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) num_cu = prop.multiProcessorCount;
    const int blocks = 4096;
    if (bf16 && single) {
      hipLaunchKernelGGL((prep_a1<__bf16, 16, true>), dim3(blocks), dim3(256), 0, s, x, n, d, dp, ntile, (__bf16*)a1);
    } else if (bf16) {
      unsigned* dmax = nullptr;
      e = hipMallocAsync((void**)&dmax, sizeof(unsigned), s);
      if (e != hipSuccess) return e;
      hipMemsetAsync(dmax, 0, sizeof(unsigned), s);
      hipLaunchKernelGGL(abs_max_kernel, dim3(2048), dim3(256), 0, s, x, n * (int64_t)d, dmax);
      unsigned hmax = 0;
      hipMemcpyAsync(&hmax, dmax, sizeof(unsigned), hipMemcpyDeviceToHost, s);
      hipFreeAsync(dmax, s);
      e = hipStreamSynchronize(s);
      if (e != hipSuccess) return e;
      float fmax_x;
      memcpy(&fmax_x, &hmax, 4);
      if (!(fmax_x < 60000.0f)) return hipErrorInvalidValue;  // fp16 GEMM2 range
      hipLaunchKernelGGL((prep_a1<__bf16, 16>), dim3(blocks), dim3(256), 0, s, x, n, d, dp, ntile, (__bf16*)a1);
      hipLaunchKernelGGL((prep_b2<_Float16, 16, true>), dim3(blocks), dim3(256), 0, s, x, n, d, dp, ntile, (_Float16*)b2);
    } else {
      hipLaunchKernelGGL((prep_a1<float, 2>), dim3(blocks), dim3(256), 0, s, x, n, d, dp, ntile, (float*)a1);
      hipLaunchKernelGGL((prep_b2<float, 2>), dim3(blocks), dim3(256), 0, s, x, n, d, dp, ntile, (float*)b2);
    }
    hipLaunchKernelGGL(pad_y, dim3(1024), dim3(256), 0, s, y, n, ntile * 32, ypad);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
  }

  hipError_t eval(hipStream_t s, const float* z, int ldz, int C, double* pe, float* g,
                  int ldg) override {
    return eval_active(s, z, ldz, C, nullptr, nullptr, pe, g, ldg);
  }
  // Only the chains act[0..*nact) (device memory) are evaluated; launch
  // geometry stays that of all C chains, tiles past *nact exit at once.
  hipError_t eval_active(hipStream_t s, const float* z, int ldz, int C, const int* act,
                         const int* nact, double* pe, float* g, int ldg) override {
    const int CT = (C + 127) / 128;
    // one 4-wave block per CU; row splits rounded to a multiple of the 8 XCDs
    int RB = std::max(8, (num_cu / CT) / 8 * 8);
    RB = (int)std::min<int64_t>(RB, (ntile + 7) / 8 * 8);
    const int nblk = CT * RB;
    // the kernels re-split the launch for the requesting chains: RB' CT' <=
    // nblk, row stride CT' 128, so nblk * 128 columns always suffice
    const size_t need = (size_t)nblk * 128 * d;
    if (need > ws_elems) {
      hipFree(wsg);
      hipFree(wsu);
      wsg = nullptr;
      wsu = nullptr;
      hipError_t e = hipMalloc(&wsg, sizeof(float) * need);
      if (e != hipSuccess) return e;
      e = hipMalloc(&wsu, sizeof(double) * (size_t)nblk * 128);
      if (e != hipSuccess) return e;
      ws_elems = need;
    }
    mark_begin(s);
#define NAX_L(FORM, DPV)                                                                      \
  hipLaunchKernelGGL((logit_mfma<FORM, DPV>), dim3(nblk), dim3(256), 0, s,                    \
                     (const FORM::frag*)a1, (const FORM::frag*)b2, ypad, n, ntile, z, ldz, d, C, \
                     wsg, wsu, act, nact)
#define NAX_DP(FORM)                    \
  switch (dp) {                         \
    case 32: NAX_L(FORM, 32); break;    \
    case 64: NAX_L(FORM, 64); break;    \
    case 96: NAX_L(FORM, 96); break;    \
    case 128: NAX_L(FORM, 128); break;  \
    case 160: NAX_L(FORM, 160); break;  \
    case 192: NAX_L(FORM, 192); break;  \
    case 224: NAX_L(FORM, 224); break;  \
    case 256: NAX_L(FORM, 256); break;  \
    default: return hipErrorInvalidValue; \
  }
#define NAX_P(DPV, SGL)                                                                        \
  hipLaunchKernelGGL((logit_bf16_pipe<DPV, SGL>), dim3(nblk), dim3(256), 0, s,                 \
                     (const bf16x8*)a1, (const f16x8*)b2, ypad, n, ntile, z, ldz, d, C, wsg,   \
                     wsu, act, nact)
#define NAX_PDP(SGL)                        \
  switch (dp) {                             \
    case 32: NAX_P(32, SGL); break;         \
    case 64: NAX_P(64, SGL); break;         \
    case 96: NAX_P(96, SGL); break;         \
    case 128: NAX_P(128, SGL); break;       \
    case 160: NAX_P(160, SGL); break;       \
    case 192: NAX_P(192, SGL); break;       \
    case 224: NAX_P(224, SGL); break;       \
    case 256: NAX_P(256, SGL); break;       \
    default: return hipErrorInvalidValue;   \
  }
    if (bf16 && single) {
      NAX_PDP(true)
    } else if (bf16) {
      NAX_PDP(false)
    } else {
      NAX_DP(FormF32)
    }
#undef NAX_PDP
#undef NAX_P
#undef NAX_DP
#undef NAX_L
    mark_end(s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const double inv_s2 = 1.0 / (prior_scale * prior_scale);
    const double log_norm = std::log(std::sqrt(2.0 * M_PI) * prior_scale);
    // padded rows each added fp32(ln 2) to every chain's energy
    const double pad_u = (double)(ntile * 32 - n) * (double)0.693147180559945309f;
    hipLaunchKernelGGL(logit_reduce, dim3(C), dim3(256), 0, s, wsg, wsu, RB, C, d, z, ldz,
                       inv_s2, log_norm, -pad_u, pe, g, ldg, act, nact, nblk, ntile);
    return hipGetLastError();
  }
};

Potential* make_logistic_mfma(const float* x, const float* y, int64_t n, int d,
                              double prior_scale, bool bf16, bool single, hipStream_t s,
                              std::string* err) {
  if (d > 256) {
    *err = "the MFMA logistic kernels support dim <= 256";
    return nullptr;
  }
  auto* p = new LogisticMfma();
  p->n = n;
  p->d = d;
  p->prior_scale = prior_scale;
  p->bf16 = bf16;
  p->single = bf16 && single;
  hipError_t e = p->setup(x, y, s);
  if (e != hipSuccess) {
    *err = std::string("logistic MFMA setup: ") + hipGetErrorString(e) +
           (bf16 && !single && e == hipErrorInvalidValue
                ? " (the two-copy bf16 form needs max|X| < 60000)" : "");
    delete p;
    return nullptr;
  }
  return p;
}

}  // namespace nax
