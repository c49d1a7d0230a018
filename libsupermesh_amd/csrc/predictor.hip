// predictor.hip -- learned output-size predictor for the single-pass tet clip.
//
// A small MLP, 24 -> 64 -> 64 -> 64 -> 33 with ReLU, maps a candidate pair
// (both tets' vertices, normalised to the pair's joint bounding box) to a
// distribution over its output count 0..32 (class 32 = "32 or more").  The
// weights are BF16 and every product accumulates in FP32 on the MFMA units
// (v_mfma_f32_32x32x16_bf16): a wave computes a 32-pair x 32-unit tile per
// instruction sequence.  Here:
//   * k_predict: features + forward + softmax; writes each pair's output
//     budget, the smallest count whose predicted cumulative probability
//     reaches the quantile q (81 = the tet clip's maximum when that is the
//     last class); activations stay in registers from layer to layer;
//   * k_train_step / k_train_wgrad / k_train_adam: one minibatch of training
//     (forward, softmax cross-entropy, backward through the ReLU layers,
//     weight gradients as MFMA reductions over the batch, Adam in FP32 on the
//     master weights, then the BF16 copies the MFMA kernels read).
// The budgets only size the output; the clip itself is the exact FP64 walk,
// and pairs whose output exceeds their budget are clipped again by the plain
// fill pass, so results never depend on the predictor.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "lsm_geometry.hpp"
#include "lsm_runtime.hpp"
#include "predictor.hpp"

namespace lsm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// BF16 weight layouts (element offsets into one array):
//   forward, [out][in] with in padded to 32 / 64: F1 [64][32], F2 [64][64],
//   F3 [64][64], F4 [64][64] (rows >= 33 zero);
//   backward, [in][out]: G2 [64][64], G3 [64][64], G4 [64][64] (cols >= 33 zero);
//   inference, [out][in] with the inputs permuted (in_pos, feat_pos) so that
//   a layer's MFMA accumulators are, lane by lane, the next layer's B
//   operand: I1 [64][32], I2 [64][64], I3 [64][64], I4 [64][64].
constexpr int kF1 = 0, kF2 = kF1 + 64 * 32, kF3 = kF2 + 64 * 64, kF4 = kF3 + 64 * 64;
constexpr int kFwdSize = kF4 + 64 * 64;  // 14336
constexpr int kG2 = kFwdSize, kG3 = kG2 + 64 * 64, kG4 = kG3 + 64 * 64;
constexpr int kI1 = kG4 + 64 * 64, kI2 = kI1 + 64 * 32, kI3 = kI2 + 64 * 64, kI4 = kI3 + 64 * 64;
constexpr int kInfSize = kFwdSize;
constexpr int kAllSize = kI4 + 64 * 64;

// A 32 x 32 MFMA accumulator holds, in lane (column c, half h = lane / 32),
// element r (0..15) at row acc_row(r) = 8 (r / 4) + 4 h + r % 4.  A B operand
// supplies, in lane (column c, half h), the 8 consecutive k of its k-step
// from 8 h.  Hidden unit n (tile t = n / 32 of a layer's output) sits in the
// accumulator of its lane's half at r = 4 ((n % 32) / 8) + n % 4; the next
// layer takes it as input position in_pos(n) = 16 s + 8 h + j with
// s = 2 t + r / 8, j = r % 8, so lane (c, h) packs accumulator elements
// 8 (s % 2) .. + 7 of tile s / 2 as its k-step s.
__host__ __device__ constexpr int in_pos(int n) {
  return 16 * (2 * (n >> 5) + ((4 * ((n & 31) >> 3) + (n & 3)) >> 3)) + 8 * (((n & 31) >> 2) & 1) +
         ((4 * ((n & 31) >> 3) + (n & 3)) & 7);
}
// Features 0..11 (A's vertices) and 12..23 (B's) at the positions the
// feature lanes supply: lane half 0 A's coordinates 0-7 (k-step 0) and 8-11
// (k-step 1), half 1 B's alike; positions 20-23 and 28-31 are zero.
__host__ __device__ constexpr int feat_pos(int f) {
  return f < 8 ? f : (f < 12 ? 16 + (f - 8) : (f < 20 ? 8 + (f - 12) : 24 + (f - 20)));
}
constexpr int kLd = 64;  // row stride (elements) of activation tiles in LDS

__device__ __forceinline__ int lane64() { return threadIdx.x & 63; }

// The value lane l ^ 32 holds (the other half of the wave), by a VALU swap
// of the halves instead of an LDS permute.
__device__ __forceinline__ float other_half(float v) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(lane64() < 32 ? r[1] : r[0]);
}

__device__ __forceinline__ double other_half(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const float lo = other_half(__uint_as_float((unsigned)b)), hi = other_half(__uint_as_float((unsigned)(b >> 32)));
  return __longlong_as_double((long long)(((unsigned long long)__float_as_uint(hi) << 32) | __float_as_uint(lo)));
}

// Row of accumulator element r in a 32x32 MFMA result (column = lane & 31).
__device__ __forceinline__ int acc_row(int r) { return 8 * (r >> 2) + 4 * (lane64() >> 5) + (r & 3); }

// acc (32 x 32) += A (32 x K) * B (K x 32).  A: rows at a + i * lda;
// B: given by columns, column j at bt + j * ldb, contiguous in k.
__device__ __forceinline__ f32x16 mma_tile(const __bf16* a, int lda, const __bf16* bt, int ldb, int K,
                                           f32x16 acc) {
  const int i = lane64() & 31, h = (lane64() >> 5) * 8;
  for (int k0 = 0; k0 < K; k0 += 16) {
    const bf16x8 va = *reinterpret_cast<const bf16x8*>(a + i * lda + k0 + h);
    const bf16x8 vb = *reinterpret_cast<const bf16x8*>(bt + i * ldb + k0 + h);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(va, vb, acc, 0, 0, 0);
  }
  return acc;
}

// Y (32 x 64, bf16) = act(X (32 x K) W^T + b); W in the forward layout.
template <int K, bool RELU>
__device__ __forceinline__ void dense(const __bf16* X, const __bf16* W, const float* bias, __bf16* Y) {
#pragma unroll
  for (int t = 0; t < 2; t++) {
    f32x16 acc = {};
    acc = mma_tile(X, kLd, W + 32 * t * K, K, K, acc);
    const int j = 32 * t + (lane64() & 31);
    const float bj = bias[j];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      float v = acc[r] + bj;
      if (RELU) v = v > 0.f ? v : 0.f;
      Y[acc_row(r) * kLd + j] = (__bf16)v;
    }
  }
}

// Features of pair (a, b): the 8 vertices normalised to their joint box.
__device__ __forceinline__ void pair_features(const double* __restrict__ ta, const double* __restrict__ tb,
                                              __bf16* row) {
  double lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = fmin_f(fmin_f(ta[d], ta[3 + d]), fmin_f(ta[6 + d], ta[9 + d]));
    hi[d] = fmax_f(fmax_f(ta[d], ta[3 + d]), fmax_f(ta[6 + d], ta[9 + d]));
    lo[d] = fmin_f(lo[d], fmin_f(fmin_f(tb[d], tb[3 + d]), fmin_f(tb[6 + d], tb[9 + d])));
    hi[d] = fmax_f(hi[d], fmax_f(fmax_f(tb[d], tb[3 + d]), fmax_f(tb[6 + d], tb[9 + d])));
  }
#pragma unroll
  for (int k = 0; k < 12; k++) {
    const int d = k % 3;
    const double s = hi[d] > lo[d] ? 1.0 / (hi[d] - lo[d]) : 0.0;
    row[k] = (__bf16)(float)((ta[k] - lo[d]) * s);
    row[12 + k] = (__bf16)(float)((tb[k] - lo[d]) * s);
  }
#pragma unroll
  for (int k = 24; k < 32; k++) row[k] = (__bf16)0.f;
}

// The quantile asked for and three below it, highest first: the single pass
// takes the highest whose runs fit in memory.  Below about 0.8 so many runs
// outgrow their budget that the spill reserve overflows and every pair is
// filled again (tet 96^3 vs 90^3: 0.95 126 ms, 0.9 130 ms, 0.85 132 ms, 0.8
// and 0.7 216 ms, plain count-then-fill 178 ms).
void budget_levels(float q, float* levels) {
  levels[0] = q;
  levels[1] = q < 0.95f ? q : 0.95f;
  levels[2] = q < 0.9f ? q : 0.9f;
  levels[3] = q < 0.8f ? q : 0.8f;
}

// ---------------------------------------------------------------------------
// Inference
// ---------------------------------------------------------------------------

constexpr int kPredWaves = 4;
constexpr int kPad = 8;  // LDS row padding (bf16) of the inference weights: rows 16 B apart mod 128

// B operand of a k-step from accumulator elements rb .. rb + 7 of tile acc:
// bias (by the hidden unit each element holds) + ReLU, rounded to BF16.
__device__ __forceinline__ bf16x8 relu_pack(const f32x16& acc, int rb, const float* bias) {
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const float x = acc[rb + j] + bias[acc_row(rb + j)];
    v[j] = (__bf16)(x > 0.f ? x : 0.f);
  }
  return v;
}

// acc[t] (units 32 t .. 32 t + 31 x 32 pairs) = W (inference layout, rows
// of K + kPad) x the 32 pairs' inputs, given as the B operands of the K / 16
// k-steps.
template <int K>
__device__ __forceinline__ void layer(const __bf16* W, const bf16x8* in, f32x16* acc) {
  const int i = lane64() & 31, h = (lane64() >> 5) * 8;
#pragma unroll
  for (int t = 0; t < 2; t++) {
    f32x16 a = {};
#pragma unroll
    for (int s = 0; s < K / 16; s++) {
      const bf16x8 va = *reinterpret_cast<const bf16x8*>(W + (32 * t + i) * (K + kPad) + 16 * s + h);
      a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(va, in[s], a, 0, 0, 0);
    }
    acc[t] = a;
  }
}

// A wave takes 32 work items at a time with no block barrier: lanes c and
// c + 32 load pair c's two tets (A, B) and build its features, the four
// layers run on the MFMA units with every activation kept in registers (the
// inference layouts put each accumulator where the next layer's operand
// wants it), and both lanes of a pair evaluate the softmax and budgets.
template <bool WORK>
__global__ void __launch_bounds__(64 * kPredWaves) __attribute__((amdgpu_waves_per_eu(4))) k_predict(PredictArgs g) {
  __shared__ __align__(16) __bf16 sW[64 * (32 + kPad) + 3 * 64 * (64 + kPad)];
  __shared__ float sB[4 * 64];
  constexpr int o1 = 0, o2 = 64 * (32 + kPad), o3 = o2 + 64 * (64 + kPad), o4 = o3 + 64 * (64 + kPad);
  for (int x = threadIdx.x; x < 64 * 32; x += blockDim.x) sW[o1 + (x >> 5) * (32 + kPad) + (x & 31)] = g.weights[kI1 + x];
  for (int x = threadIdx.x; x < 64 * 64; x += blockDim.x) {
    const int r = x >> 6, c = x & 63;
    sW[o2 + r * (64 + kPad) + c] = g.weights[kI2 + x];
    sW[o3 + r * (64 + kPad) + c] = g.weights[kI3 + x];
    sW[o4 + r * (64 + kPad) + c] = g.weights[kI4 + x];
  }
  for (int x = threadIdx.x; x < 4 * 64; x += blockDim.x) sB[x] = g.bias[x];
  __syncthreads();
  const int l = lane64(), c = l & 31, half = l >> 5;
  const long ntiles = (g.nwork + 31) / 32;
  const long nwaves = (long)gridDim.x * kPredWaves;
  unsigned long long tot[5] = {0, 0, 0, 0, 0};
  // the pair of this lane's work item in tile `tile` (past the end, that of
  // the last item: a lane past the end computes on real data, and its pair's
  // column of the MFMA tiles is never stored), and the element of this
  // lane's tet of pair p (A's in half 0, B's in half 1).  Three tiles are in
  // flight: each tile issues the work item load of the tile two ahead, the
  // element load of the next tile and its own coordinate loads back to back
  // and unconditionally, so that it waits on one round trip (loads complete
  // in order: a wait for any of them waits for the earlier ones)
  auto pair_of = [&](long tile) -> int {
    const long w = std::min<long>(tile * 32 + c, g.nwork - 1);
    return WORK ? g.work[w] : (int)w;
  };
  const int* element = half ? g.cols : g.rows;
  long tile = (long)blockIdx.x * kPredWaves + (threadIdx.x >> 6);
  int e = element[pair_of(tile)];
  int p_next = pair_of(tile + nwaves);
  // waited for here, before the loop, so that at the loop head only the
  // previous tile's budget store can be in flight, and no load waits for it
  asm volatile("" : "+v"(e), "+v"(p_next));
  for (; tile < ntiles; tile += nwaves) {
    // the weights are read from LDS at each tile: held in registers across
    // tiles (where the compiler would hoist them) they would not fit
    int zero = 0;
    asm volatile("" : "+v"(zero));
    const __bf16* W = sW + zero;
    const long w = tile * 32 + c;
    const bool valid = w < g.nwork;
    const int p_after = pair_of(tile + 2 * nwaves);
    const int e_next = element[p_next];
    double T[12];
    {
      const double2* src = reinterpret_cast<const double2*>((half ? g.bcoords : g.acoords) + (long)e * 12);
#pragma unroll
      for (int k = 0; k < 6; k++) {
        const double2 v = src[k];
        T[2 * k] = v.x;
        T[2 * k + 1] = v.y;
      }
    }
    e = e_next;
    p_next = p_after;
    // features: both tets normalised to their joint box (pair_features)
    bf16x8 x[2];
    {
      double lo[3], hi[3], sc[3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        lo[d] = fmin_f(fmin_f(T[d], T[3 + d]), fmin_f(T[6 + d], T[9 + d]));
        hi[d] = fmax_f(fmax_f(T[d], T[3 + d]), fmax_f(T[6 + d], T[9 + d]));
        lo[d] = fmin_f(lo[d], other_half(lo[d]));
        hi[d] = fmax_f(hi[d], other_half(hi[d]));
        sc[d] = hi[d] > lo[d] ? 1.0 / (hi[d] - lo[d]) : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 8; k++) x[0][k] = (__bf16)(float)((T[k] - lo[k % 3]) * sc[k % 3]);
#pragma unroll
      for (int k = 0; k < 8; k++)
        x[1][k] = k < 4 ? (__bf16)(float)((T[8 + k] - lo[(8 + k) % 3]) * sc[(8 + k) % 3]) : (__bf16)0.f;
    }
    f32x16 acc[2];
    bf16x8 b[4];
    layer<32>(W + o1, x, acc);
#pragma unroll
    for (int s = 0; s < 4; s++) b[s] = relu_pack(acc[s >> 1], 8 * (s & 1), sB + 32 * (s >> 1));
    layer<64>(W + o2, b, acc);
#pragma unroll
    for (int s = 0; s < 4; s++) b[s] = relu_pack(acc[s >> 1], 8 * (s & 1), sB + 64 + 32 * (s >> 1));
    layer<64>(W + o3, b, acc);
#pragma unroll
    for (int s = 0; s < 4; s++) b[s] = relu_pack(acc[s >> 1], 8 * (s & 1), sB + 128 + 32 * (s >> 1));
    layer<64>(W + o4, b, acc);
    // logits of pair c: this lane's 16 classes 8 (r / 4) + 4 half + r % 4,
    // and unit 32 (tile 1, row 0) in half 0.  The softmax, its running sums
    // and the four budget levels are split between the pair's two lanes:
    // each exponentiates its own classes, and the halves swap their maximum,
    // sums, the sums of their four 4-class blocks and their level counts.
    float v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) v[r] = acc[0][r] + sB[192 + 8 * (r >> 2) + 4 * half + (r & 3)];
    const float v32 = acc[1][0] + sB[192 + 32];
    float mloc = half == 0 ? v32 : v[0];
#pragma unroll
    for (int r = 0; r < 16; r++) mloc = fmaxf(mloc, v[r]);
    const float mx = fmaxf(mloc, other_half(mloc));
    float e[16], bs[4];
    float sloc = 0.f, exloc = 0.f;
#pragma unroll
    for (int jb = 0; jb < 4; jb++) {
      bs[jb] = 0.f;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int r = 4 * jb + i;
        e[r] = __expf(v[r] - mx);
        bs[jb] += e[r];
        exloc += (float)(8 * jb + 4 * half + i) * e[r];
      }
      sloc += bs[jb];
    }
    const float e32 = half == 0 ? __expf(v32 - mx) : 0.f;
    sloc += e32;
    exloc += 32.f * e32;
    const float ssum = sloc + other_half(sloc), exsum = exloc + other_half(exloc);
    // probability before each of this lane's blocks: block 2 jb + h (classes
    // 8 jb + 4 h .. + 3) is half h's block jb
    float obs[4];  // the other half's block sums (swapped by every lane alike)
#pragma unroll
    for (int jb = 0; jb < 4; jb++) obs[jb] = other_half(bs[jb]);
    float before[4] = {0.f, 0.f, 0.f, 0.f};
    float run = 0.f;
#pragma unroll
    for (int B = 0; B < 8; B++) {
      const bool mine = (B & 1) == half;
      if (mine) before[B >> 1] = run;
      run += mine ? bs[B >> 1] : obs[B >> 1];
    }
    float thr[4];
#pragma unroll
    for (int lv = 0; lv < 4; lv++) thr[lv] = g.levels[lv] * ssum;
    // classes below 32 whose cumulative probability stays under each level,
    // one byte per level (at most 16 per lane)
    unsigned cnt = 0;
#pragma unroll
    for (int jb = 0; jb < 4; jb++) {
      float cum = before[jb];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        cum += e[4 * jb + i];
#pragma unroll
        for (int lv = 0; lv < 4; lv++) cnt += cum < thr[lv] ? (1u << (8 * lv)) : 0u;
      }
    }
    cnt += __float_as_uint(other_half(__uint_as_float(cnt)));
    unsigned bud = 0;
#pragma unroll
    for (int lv = 0; lv < 4; lv++) {
      const unsigned n = (cnt >> (8 * lv)) & 0xffu;
      bud |= (n == 32 ? (unsigned)kTetBufSize : n) << (8 * lv);
    }
    // every lane stores (both lanes of a pair hold its budget; past the end,
    // to the spare word), so that the store is counted exactly and the next
    // tile's loads do not wait for it to complete
    *(valid ? g.budget + w : g.spare) = bud;
    if (valid && half == 0) {
#pragma unroll
      for (int k = 0; k < 4; k++) tot[k] += (bud >> (8 * k)) & 0xffu;
      tot[4] += (unsigned)(1024.f * exsum / ssum);
    }
  }
#pragma unroll
  for (int k = 0; k < 5; k++) {
    unsigned long long v = tot[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (l == 0 && v) atomicAdd(g.sums + k, v);
  }
}

// ---------------------------------------------------------------------------
// Training
// ---------------------------------------------------------------------------

constexpr int kTrainWaves = 2;

// Forward and backward of one 32-sample tile per wave.  Writes the transposed
// activations HT_l [64][B] (l = 0: the features, [32][B]) and output-side
// gradients DT_l [64][B] (l = 1..4) that the weight-gradient kernel reduces.
__global__ void __launch_bounds__(64 * kTrainWaves) k_train_step(TrainArgs g) {
  __shared__ __align__(16) __bf16 sH[kTrainWaves][4][32 * kLd];  // H1 H2 H3 D
  __shared__ float sL[kTrainWaves][32 * 33];
  const int wv = threadIdx.x >> 6, l = lane64();
  // the batch is a multiple of 32 * kTrainWaves: every wave has a tile
  const long tile = (long)blockIdx.x * kTrainWaves + wv;
  const long base = tile * 32;
  __bf16* H1 = sH[wv][0];
  __bf16* H2 = sH[wv][1];
  __bf16* H3 = sH[wv][2];
  __bf16* D = sH[wv][3];
  float* L = sL[wv];
  const __bf16* W = g.weights;
  const long B = g.batch;
  // features of the samples -> D (as X0) and HT_0
  if (l < 32) {
    const long s = g.index[base + l];
    for (int k = 0; k < 32; k++) {
      const __bf16 v = g.features[s * 32 + k];
      D[l * kLd + k] = v;
      g.ht0[(long)k * B + base + l] = v;
    }
  }
  __syncthreads();
  dense<32, true>(D, W + kF1, g.bias, H1);
  __syncthreads();
  dense<64, true>(H1, W + kF2, g.bias + 64, H2);
  __syncthreads();
  dense<64, true>(H2, W + kF3, g.bias + 128, H3);
  __syncthreads();
  f32x16 a0 = {}, a1 = {};
  a0 = mma_tile(H3, kLd, W + kF4, 64, 64, a0);
  a1 = mma_tile(H3, kLd, W + kF4 + 32 * 64, 64, 64, a1);
  {
    const int j = l & 31;
#pragma unroll
    for (int r = 0; r < 16; r++) L[acc_row(r) * 33 + j] = a0[r] + g.bias[192 + j];
    if (j == 0) {
#pragma unroll
      for (int r = 0; r < 16; r++) L[acc_row(r) * 33 + 32] = a1[r] + g.bias[192 + 32];
    }
  }
  // transposed activations for the weight gradients
  for (int k = 0; k < 64; k++) {
    if (l < 32) {
      g.ht1[(long)k * B + base + l] = H1[l * kLd + k];
      g.ht2[(long)k * B + base + l] = H2[l * kLd + k];
      g.ht3[(long)k * B + base + l] = H3[l * kLd + k];
    }
  }
  __syncthreads();
  // softmax cross-entropy: D = dL/dlogits = (p - onehot) / B
  if (l < 32) {
    const float* row = L + l * 33;
    const int y = g.labels[g.index[base + l]];
    float m = row[0];
    for (int c = 1; c < 33; c++) m = fmaxf(m, row[c]);
    float s = 0.f;
    for (int c = 0; c < 33; c++) s += __expf(row[c] - m);
    const float inv = 1.f / s;
    for (int c = 0; c < 64; c++) {
      float d = 0.f;
      if (c < 33) d = (__expf(row[c] - m) * inv - (c == y ? 1.f : 0.f)) / (float)B;
      const __bf16 v = (__bf16)d;
      D[l * kLd + c] = v;
      g.dt4[(long)c * B + base + l] = v;
    }
    atomicAdd(g.loss, (m + __logf(s)) - row[y]);
  }
  __syncthreads();
  // back through layers 4, 3, 2: dZ_{l} = (dZ_{l+1} W_{l+1}) * (H_l > 0)
  const __bf16* Gs[3] = {W + kG4, W + kG3, W + kG2};
  __bf16* Hs[3] = {H3, H2, H1};
  __bf16* DT[3] = {g.dt3, g.dt2, g.dt1};
  for (int s = 0; s < 3; s++) {
    f32x16 c0 = {}, c1 = {};
    c0 = mma_tile(D, kLd, Gs[s], 64, 64, c0);
    c1 = mma_tile(D, kLd, Gs[s] + 32 * 64, 64, 64, c1);
    __syncthreads();
    const int j = l & 31;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int i = acc_row(r);
      const float h0 = (float)Hs[s][i * kLd + j], h1 = (float)Hs[s][i * kLd + 32 + j];
      D[i * kLd + j] = (__bf16)(h0 > 0.f ? c0[r] : 0.f);
      D[i * kLd + 32 + j] = (__bf16)(h1 > 0.f ? c1[r] : 0.f);
    }
    __syncthreads();
    for (int k = 0; k < 64; k++)
      if (l < 32) DT[s][(long)k * B + base + l] = D[l * kLd + k];
    __syncthreads();
  }
}

// Weight gradients: grad_l [out][in] = sum_k DT_l[out][k] HT_{l-1}[in][k],
// one wave per 32 x 32 tile; bias gradients from the first column tile.
__global__ void __launch_bounds__(64) k_train_wgrad(TrainArgs g) {
  // tile list: layer 1: 2 (out 64 x in 32), layers 2..4: 4 each
  int t = blockIdx.x, layer, mt, nt;
  if (t < 2) {
    layer = 0;
    mt = t;
    nt = 0;
  } else {
    t -= 2;
    layer = 1 + t / 4;
    mt = (t % 4) / 2;
    nt = t % 2;
  }
  const __bf16* dts[4] = {g.dt1, g.dt2, g.dt3, g.dt4};
  const __bf16* hts[4] = {g.ht0, g.ht1, g.ht2, g.ht3};
  const int nin = layer == 0 ? 24 : 64, nout = layer == 3 ? 33 : 64;
  const int woff[4] = {kOffW1, kOffW2, kOffW3, kOffW4}, boff[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
  const long B = g.batch;
  const __bf16* A = dts[layer] + (long)(32 * mt) * B;
  const __bf16* Bt = hts[layer] + (long)(32 * nt) * B;
  f32x16 acc = {};
  acc = mma_tile(A, (int)B, Bt, (int)B, (int)B, acc);
  const int j = 32 * nt + (lane64() & 31);
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int o = 32 * mt + acc_row(r);
    if (o < nout && j < nin) g.grad[woff[layer] + o * nin + j] = acc[r];
  }
  if (nt == 0 && lane64() < 32) {
    const int o = 32 * mt + lane64();
    if (o < nout) {
      float s = 0.f;
      const __bf16* row = dts[layer] + (long)o * B;
      for (long k = 0; k < B; k++) s += (float)row[k];
      g.grad[boff[layer] + o] = s;
    }
  }
}

// Writes the BF16 forward / backward copies and biases of parameter i.
__device__ __forceinline__ void publish(int i, float w, __bf16* W, float* bias) {
  const __bf16 v = (__bf16)w;
  if (i < kOffB1) {
    const int o = (i - kOffW1) / 24, n = (i - kOffW1) % 24;
    W[kF1 + o * 32 + n] = v;
    W[kI1 + o * 32 + feat_pos(n)] = v;
  } else if (i < kOffW2) {
    bias[i - kOffB1] = w;
  } else if (i < kOffB2) {
    const int o = (i - kOffW2) / 64, n = (i - kOffW2) % 64;
    W[kF2 + o * 64 + n] = v;
    W[kG2 + n * 64 + o] = v;
    W[kI2 + o * 64 + in_pos(n)] = v;
  } else if (i < kOffW3) {
    bias[64 + i - kOffB2] = w;
  } else if (i < kOffB3) {
    const int o = (i - kOffW3) / 64, n = (i - kOffW3) % 64;
    W[kF3 + o * 64 + n] = v;
    W[kG3 + n * 64 + o] = v;
    W[kI3 + o * 64 + in_pos(n)] = v;
  } else if (i < kOffW4) {
    bias[128 + i - kOffB3] = w;
  } else if (i < kOffB4) {
    const int o = (i - kOffW4) / 64, n = (i - kOffW4) % 64;
    W[kF4 + o * 64 + n] = v;
    W[kG4 + n * 64 + o] = v;
    W[kI4 + o * 64 + in_pos(n)] = v;
  } else {
    bias[192 + i - kOffB4] = w;
  }
}

__global__ void __launch_bounds__(256) k_train_adam(TrainArgs g, float lr, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNumParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gr = g.grad[i];
  const float m = b1 * g.adam_m[i] + (1.f - b1) * gr;
  const float v = b2 * g.adam_v[i] + (1.f - b2) * gr * gr;
  g.adam_m[i] = m;
  g.adam_v[i] = v;
  const float w = g.master[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
  g.master[i] = w;
  publish(i, w, g.weights, g.bias);
}

__global__ void __launch_bounds__(256) k_publish(const float* __restrict__ master, __bf16* W, float* bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kNumParams) publish(i, master[i], W, bias);
}

__global__ void __launch_bounds__(256)
k_zero_bf16(long n, __bf16* __restrict__ p) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (__bf16)0.f;
}

// Features and labels of sampled work pairs (one thread per sample).
__global__ void __launch_bounds__(256)
k_collect(long n, long stride, const int* __restrict__ work, const int* __restrict__ rows,
          const int* __restrict__ cols, const int* __restrict__ counts, const double* __restrict__ ac,
          const double* __restrict__ bc, __bf16* __restrict__ features, int* __restrict__ labels) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const long p = work[s * stride];
  __bf16 row[32];
  pair_features(ac + (long)rows[p] * 12, bc + (long)cols[p] * 12, row);
  for (int k = 0; k < 32; k++) features[s * 32 + k] = row[k];
  const int c = counts[p];
  labels[s] = c < 32 ? c : 32;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static void ensure_weights(Context& C) {
  Predictor& P = C.predictor;
  if (P.weights.ptr) return;
  __bf16* W = P.weights.as<__bf16>(kAllSize);
  hipLaunchKernelGGL(k_zero_bf16, dim3(grid_for(kAllSize, 256)), dim3(256), 0, C.stream, (long)kAllSize, W);
  P.bias.as<float>(256);
  LSM_HIP(hipMemsetAsync(P.bias.ptr, 0, 256 * sizeof(float), C.stream));
  LSM_HIP(hipGetLastError());
}

void predictor_set(Context& C, const float* params) {
  Predictor& P = C.predictor;
  ensure_weights(C);
  float* m = P.master.as<float>(kNumParams);
  LSM_HIP(hipMemcpyAsync(m, params, sizeof(float) * kNumParams, hipMemcpyHostToDevice, C.stream));
  LSM_HIP(hipMemsetAsync(P.adam_m.as<float>(kNumParams), 0, sizeof(float) * kNumParams, C.stream));
  LSM_HIP(hipMemsetAsync(P.adam_v.as<float>(kNumParams), 0, sizeof(float) * kNumParams, C.stream));
  hipLaunchKernelGGL(k_publish, dim3(grid_for(kNumParams, 256)), dim3(256), 0, C.stream, m,
                     static_cast<__bf16*>(P.weights.ptr), static_cast<float*>(P.bias.ptr));
  LSM_HIP(hipGetLastError());
  LSM_HIP(hipStreamSynchronize(C.stream));
  P.ready = true;
  P.steps = 0;
}

void predictor_get(Context& C, float* params) {
  Predictor& P = C.predictor;
  LSM_REQUIRE(P.ready, LIBSUPERMESH_GPU_ERROR_STATE, "no predictor weights loaded");
  LSM_HIP(hipMemcpyAsync(params, P.master.ptr, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
}

void predictor_init(Context& C, unsigned seed) {
  // He-uniform weights, zero biases
  std::vector<float> p(kNumParams, 0.f);
  std::mt19937 rng(seed);
  auto fill = [&](int off, int nout, int nin) {
    const float a = std::sqrt(6.f / (float)nin);
    std::uniform_real_distribution<float> U(-a, a);
    for (int i = 0; i < nout * nin; i++) p[off + i] = U(rng);
  };
  fill(kOffW1, 64, 24);
  fill(kOffW2, 64, 64);
  fill(kOffW3, 64, 64);
  fill(kOffW4, 33, 64);
  predictor_set(C, p.data());
}

static long reserve_samples(Context& C, long extra);

void predictor_collect(Context& C, long nwork, const int* work, const int* rows, const int* cols,
                       const int* counts, const double* acoords, const double* bcoords, long max_samples) {
  Predictor& P = C.predictor;
  if (nwork <= 0 || max_samples <= 0) return;
  const long stride = (nwork + max_samples - 1) / max_samples;
  const long n = (nwork + stride - 1) / stride;
  const long old = reserve_samples(C, n);
  __bf16* f = static_cast<__bf16*>(P.features.ptr) + 32 * old;
  int* lab = static_cast<int*>(P.labels.ptr) + old;
  hipLaunchKernelGGL(k_collect, dim3(grid_for(n, 256)), dim3(256), 0, C.stream, n, stride, work, rows, cols,
                     counts, acoords, bcoords, f, lab);
  LSM_HIP(hipGetLastError());
  LSM_HIP(hipStreamSynchronize(C.stream));
  P.nsamples = old + n;
}

__global__ void __launch_bounds__(256)
k_samples_from_host(long n, const float* __restrict__ f, const int* __restrict__ y, __bf16* __restrict__ features,
                    int* __restrict__ labels) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  for (int k = 0; k < 32; k++) features[s * 32 + k] = (__bf16)(k < 24 ? f[s * 24 + k] : 0.f);
  const int c = y[s];
  labels[s] = c < 0 ? 0 : (c < 32 ? c : 32);
}

// Grows the sample store to hold `extra` more samples; returns the old count.
static long reserve_samples(Context& C, long extra) {
  Predictor& P = C.predictor;
  const long old = P.nsamples;
  const size_t fbytes = sizeof(__bf16) * 32 * (size_t)(old + extra), lbytes = sizeof(int) * (size_t)(old + extra);
  if (fbytes > P.features.bytes || lbytes > P.labels.bytes) {
    DevBuf nf, nl;
    nf.get(fbytes);
    nl.get(lbytes);
    if (old > 0) {
      LSM_HIP(hipMemcpyAsync(nf.ptr, P.features.ptr, sizeof(__bf16) * 32 * old, hipMemcpyDeviceToDevice, C.stream));
      LSM_HIP(hipMemcpyAsync(nl.ptr, P.labels.ptr, sizeof(int) * old, hipMemcpyDeviceToDevice, C.stream));
    }
    LSM_HIP(hipStreamSynchronize(C.stream));
    P.features.release();
    P.labels.release();
    P.features = nf;
    P.labels = nl;
  }
  return old;
}

void predictor_add_samples(Context& C, long n, const float* features, const int* labels) {
  if (n <= 0) return;
  Predictor& P = C.predictor;
  const long old = reserve_samples(C, n);
  DevBuf hf, hy;
  float* df = static_cast<float*>(hf.get(sizeof(float) * 24 * n));
  int* dy = static_cast<int*>(hy.get(sizeof(int) * n));
  LSM_HIP(hipMemcpyAsync(df, features, sizeof(float) * 24 * n, hipMemcpyHostToDevice, C.stream));
  LSM_HIP(hipMemcpyAsync(dy, labels, sizeof(int) * n, hipMemcpyHostToDevice, C.stream));
  hipLaunchKernelGGL(k_samples_from_host, dim3(grid_for(n, 256)), dim3(256), 0, C.stream, n, df, dy,
                     static_cast<__bf16*>(P.features.ptr) + 32 * old, static_cast<int*>(P.labels.ptr) + old);
  LSM_HIP(hipGetLastError());
  LSM_HIP(hipStreamSynchronize(C.stream));
  hf.release();
  hy.release();
  P.nsamples = old + n;
}

void predictor_train(Context& C, int epochs, int batch, float lr, unsigned seed, float* loss_out) {
  Predictor& P = C.predictor;
  if (!P.ready) predictor_init(C, seed);
  LSM_REQUIRE(batch >= 64 && batch % 64 == 0, LIBSUPERMESH_GPU_ERROR_INVALID,
              "training batch must be a positive multiple of 64");
  LSM_REQUIRE(P.nsamples >= batch, LIBSUPERMESH_GPU_ERROR_INVALID, "fewer training samples than one batch");
  const long n = P.nsamples, steps = n / batch;
  TrainArgs g = {};
  g.batch = batch;
  g.features = static_cast<const __bf16*>(P.features.ptr);
  g.labels = static_cast<const int*>(P.labels.ptr);
  g.weights = static_cast<__bf16*>(P.weights.ptr);
  g.bias = static_cast<float*>(P.bias.ptr);
  g.master = static_cast<float*>(P.master.ptr);
  g.adam_m = static_cast<float*>(P.adam_m.ptr);
  g.adam_v = static_cast<float*>(P.adam_v.ptr);
  g.grad = P.grad.as<float>(kNumParams);
  LSM_HIP(hipMemsetAsync(g.grad, 0, sizeof(float) * kNumParams, C.stream));
  __bf16* act = P.act.as<__bf16>((size_t)batch * (32 + 64 * 3 + 64 * 4));
  g.ht0 = act;
  g.ht1 = g.ht0 + (size_t)32 * batch;
  g.ht2 = g.ht1 + (size_t)64 * batch;
  g.ht3 = g.ht2 + (size_t)64 * batch;
  g.dt1 = g.ht3 + (size_t)64 * batch;
  g.dt2 = g.dt1 + (size_t)64 * batch;
  g.dt3 = g.dt2 + (size_t)64 * batch;
  g.dt4 = g.dt3 + (size_t)64 * batch;
  float* dloss = P.loss.as<float>(1);
  int* didx = P.index.as<int>((size_t)n);
  std::vector<int> perm((size_t)n);
  for (long i = 0; i < n; i++) perm[i] = (int)i;
  std::mt19937 rng(seed * 2654435761u + 1u);
  const float b1 = 0.9f, b2 = 0.999f;
  for (int e = 0; e < epochs; e++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    LSM_HIP(hipMemcpyAsync(didx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, C.stream));
    LSM_HIP(hipMemsetAsync(dloss, 0, sizeof(float), C.stream));
    g.loss = dloss;
    for (long s = 0; s < steps; s++) {
      g.index = didx + s * batch;
      hipLaunchKernelGGL(k_train_step, dim3((unsigned)(batch / (32 * kTrainWaves))),
                         dim3(64 * kTrainWaves), 0, C.stream, g);
      hipLaunchKernelGGL(k_train_wgrad, dim3(14), dim3(64), 0, C.stream, g);
      P.steps++;
      const float c1 = 1.f - std::pow(b1, (float)P.steps), c2 = 1.f - std::pow(b2, (float)P.steps);
      hipLaunchKernelGGL(k_train_adam, dim3(grid_for(kNumParams, 256)), dim3(256), 0, C.stream, g, lr, c1, c2);
    }
    LSM_HIP(hipGetLastError());
    float h = 0.f;
    LSM_HIP(hipMemcpyAsync(&h, dloss, sizeof(float), hipMemcpyDeviceToHost, C.stream));
    LSM_HIP(hipStreamSynchronize(C.stream));
    if (loss_out) loss_out[e] = steps > 0 ? h / (float)(steps * batch) : 0.f;
  }
  // a held copy of the perm vector is not needed beyond this call
}

void predict_budgets(Context& C, const PredictArgs& args) {
  if (args.nwork <= 0) return;
  int cus = 0;
  LSM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, C.device));
  const long tiles = (args.nwork + 31) / 32;
  long blocks = std::min<long>((long)(cus > 0 ? cus : 256) * 4, (tiles + kPredWaves - 1) / kPredWaves);
  blocks = std::max<long>(blocks, 1);
  if (args.work)
    hipLaunchKernelGGL(k_predict<true>, dim3((unsigned)blocks), dim3(64 * kPredWaves), 0, C.stream, args);
  else
    hipLaunchKernelGGL(k_predict<false>, dim3((unsigned)blocks), dim3(64 * kPredWaves), 0, C.stream, args);
  LSM_HIP(hipGetLastError());
}

}  // namespace lsm
