// cf_mlp.hpp -- the learned SGS closure network on the CDNA4 matrix cores.
//
// A pointwise MLP 11 -> 64 -> 64 -> 64 -> 6 with GELU (tanh form) after each
// hidden layer.  No counterpart in the reference: it replaces the stresses of
// compsgs1.f90:1-53 (see k_rows_mlp).  BF16 operands on
// v_mfma_f32_32x32x16_bf16 with FP32 accumulation, for inference, the
// training forward and backward passes; Adam on FP32 master weights.
//
// Wavefront layout.  A wave evaluates 64 samples (lane = sample) as two
// 32-sample tiles.  Activations are kept in the MFMA accumulator layout
// ("D layout"): for a [units x 32 samples] tile mt, lane l and accumulator i
// hold unit d_unit(mt, i, l>>5) = 32mt + 8(i>>2) + 4(l>>5) + (i&3) of sample
// l&31.  Accumulators 8(s&1)..8(s&1)+7 of tile s>>1, converted to BF16, are
// exactly the B operand of K-step s of the next layer once the next layer's
// weights are stored with their input units permuted by chain_unit(); no
// lane exchange happens between layers.  Only the K index pairing of A and B
// has to agree, and both tables are built with the same (lane, element) -> k
// map, so the result does not depend on how the hardware splits K.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cf {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int NIN = 11, NH = 64, NOUT = 6;
// flat parameters: W1 [64][11], b1, W2 [64][64], b2, W3 [64][64], b3, W4 [6][64], b4
constexpr int OW1 = 0, OB1 = 704, OW2 = 768, OB2 = 4864, OW3 = 4928, OB3 = 9024, OW4 = 9088,
              OB4 = 9472, NPARAM = 9478;
// forward A-operand tiles (64 lanes x 8 BF16 each): layer 1 mt=0..1; layers
// 2, 3 tile mt*4+s; layer 4 tile s
constexpr int FT1 = 0, FT2 = 2, FT3 = 10, FT4 = 18, FTN = 22;
// backward A-operand tiles (transposed weights): layer 4 mt=0..1; layers 3, 2 mt*4+s
constexpr int BT4 = 0, BT3 = 2, BT2 = 10, BTN = 18;
// bias tables in D layout: layers 1..3 [mt][half][16], layer 4 [half][16]
constexpr int FB1 = 0, FB2 = 64, FB3 = 128, FB4 = 192, FBN = 224;

// A published network: everything inference needs, one 16-byte aligned blob.
struct Net {
  bf16x8 fw[FTN * 64];
  float fb[FBN];
  float in_mu[16];      // x = (f - in_mu) * in_rs for the 11 inputs
  float in_rs[16];
  float out_sd[8];      // tau = y * out_sd + out_mu
  float out_mu[8];
};
static_assert(sizeof(Net) % 16 == 0, "Net is copied in 16-byte words");

__host__ __device__ constexpr int chain_unit(int s, int half, int j) {
  return 32 * (s >> 1) + 16 * (s & 1) + 8 * (j >> 2) + 4 * half + (j & 3);
}
__host__ __device__ constexpr int d_unit(int mt, int i, int half) {
  return 32 * mt + 8 * (i >> 2) + 4 * half + (i & 3);
}

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// GELU, tanh form: 0.5 z (1 + tanh(k (z + a z^3))) = z / (1 + 2^(c1 z + c2 z^3))
// with k = sqrt(2/pi), a = 0.044715, c1 = -2 k log2(e), c2 = a c1.
constexpr float GELU_C1 = -2.30220819f, GELU_C2 = -0.102943241f;
__device__ __forceinline__ float gelu_sig(float z) {
  const float t = z * __builtin_fmaf(GELU_C2, z * z, GELU_C1);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_sig(z); }
__device__ __forceinline__ float gelu_grad(float z) {
  const float s = gelu_sig(z);
  return s + z * s * (1.0f - s) * 1.59576912f * __builtin_fmaf(0.134145f, z * z, 1.0f);
}

__device__ __forceinline__ bf16x8 pack8(const f32x16& d, int base) {
  bf16x8 b;
#pragma unroll
  for (int j = 0; j < 8; ++j) b[j] = (__bf16)d[base + j];
  return b;
}
__device__ __forceinline__ bf16x8 pack8_gelu(const f32x16& d, int base) {
  bf16x8 b;
#pragma unroll
  for (int j = 0; j < 8; ++j) b[j] = (__bf16)gelu(d[base + j]);
  return b;
}
__device__ __forceinline__ f32x16 load16(const float* p) {
  f32x16 v;
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = p[i];
  return v;
}
__device__ __forceinline__ bf16x8 shfl_xor32(bf16x8 v) {
  u32x4 u = __builtin_bit_cast(u32x4, v);
#pragma unroll
  for (int q = 0; q < 4; ++q) u[q] = (unsigned)__shfl_xor((int)u[q], 32, 64);
  return __builtin_bit_cast(bf16x8, u);
}

// Keeps the scheduler from hoisting one layer's weight loads into another's
// (each hoisted A operand costs 4 VGPRs in the row kernel).
__device__ __forceinline__ void layer_fence() { asm volatile("" ::: "memory"); }

// Hidden layer with K-steps S: tile mt of b + W B, then its GELU packed into
// the next layer's operands Bn[2mt], Bn[2mt+1].  Tile 1's MFMAs overlap tile
// 0's GELU, and only one accumulator tile is live.
template <int S, bool SAVE>
__device__ __forceinline__ void hidden(const Net* net, int ft, int fb, const bf16x8 (&B)[4],
                                       bf16x8 (&Bn)[4], int lane, float* zs, int layer) {
  const int half = lane >> 5;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    layer_fence();
    f32x16 D = load16(net->fb + fb + (mt * 2 + half) * 16);
#pragma unroll
    for (int s = 0; s < S; ++s) D = mfma(net->fw[(ft + mt * S + s) * 64 + lane], B[s], D);
    if constexpr (SAVE) {
#pragma unroll
      for (int i = 0; i < 16; ++i) zs[((layer * 2 + mt) * 16 + i) * 64 + lane] = D[i];
    }
    Bn[2 * mt] = pack8_gelu(D, 0);
    Bn[2 * mt + 1] = pack8_gelu(D, 8);
  }
}

// The four layers on one 32-sample tile whose input B operand is xin.  With
// SAVE, the pre-activations of layers 1..3 go to zs[((layer*2+mt)*16+i)*64+lane].
template <bool SAVE>
__device__ __forceinline__ f32x16 forward_tile(const Net* net, bf16x8 xin, int lane, float* zs) {
  const int half = lane >> 5;
  bf16x8 B[4], C[4];
  B[0] = xin;
  hidden<1, SAVE>(net, FT1, FB1, B, C, lane, zs, 0);
  hidden<4, SAVE>(net, FT2, FB2, C, B, lane, zs, 1);
  hidden<4, SAVE>(net, FT3, FB3, B, C, lane, zs, 2);
  layer_fence();
  f32x16 Y = load16(net->fb + FB4 + half * 16);
#pragma unroll
  for (int s = 0; s < 4; ++s) Y = mfma(net->fw[(FT4 + s) * 64 + lane], C[s], Y);
  return Y;
}

// Normalised features of this lane's sample -> the input B operands of the
// wave's two 32-sample tiles (lanes < 32 carry inputs 0..7, lanes >= 32
// inputs 8..15 of sample lane&31 + 32 nt).
__device__ __forceinline__ void input_operands(const float (&x)[16], int lane, bf16x8& b0,
                                               bf16x8& b1) {
  bf16x8 lo, hi;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    lo[j] = (__bf16)x[j];
    hi[j] = (__bf16)x[8 + j];
  }
  const bool upper = lane >= 32;
  const bf16x8 recv = shfl_xor32(upper ? lo : hi);
  b0 = upper ? recv : lo;
  b1 = upper ? hi : recv;
}

// Output tiles of the wave -> this lane's 6 outputs (its own sample).
__device__ __forceinline__ void output_of_lane(const f32x16& y0, const f32x16& y1, int lane,
                                               float (&y)[NOUT]) {
  const bool upper = lane >= 32;
  float s[4] = {upper ? y0[0] : y1[0], upper ? y0[1] : y1[1], upper ? 0.f : y1[2],
                upper ? 0.f : y1[3]};
  float r[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) r[q] = __shfl_xor(s[q], 32, 64);
  if (!upper) {
    y[0] = y0[0]; y[1] = y0[1]; y[2] = y0[2]; y[3] = y0[3]; y[4] = r[0]; y[5] = r[1];
  } else {
    y[0] = r[0]; y[1] = r[1]; y[2] = r[2]; y[3] = r[3]; y[4] = y1[0]; y[5] = y1[1];
  }
}

// Inference for the 64 samples of a wave: raw features f of this lane's
// sample -> its 6 raw stresses.  Every lane of the wave must call it.
__device__ __forceinline__ void infer(const Net* net, const float (&f)[NIN], float (&tau)[NOUT],
                                      int lane) {
  float x[16];
#pragma unroll
  for (int k = 0; k < NIN; ++k) x[k] = (f[k] - net->in_mu[k]) * net->in_rs[k];
#pragma unroll
  for (int k = NIN; k < 16; ++k) x[k] = 0.f;
  bf16x8 b0, b1;
  input_operands(x, lane, b0, b1);
  const f32x16 y0 = forward_tile<false>(net, b0, lane, nullptr);
  const f32x16 y1 = forward_tile<false>(net, b1, lane, nullptr);
  float y[NOUT];
  output_of_lane(y0, y1, lane, y);
#pragma unroll
  for (int k = 0; k < NOUT; ++k) tau[k] = y[k] * net->out_sd[k] + net->out_mu[k];
}

// Realizability (non-negative principal minors) and no backscatter
// (-tau_ij S_ij >= 0) of a predicted stress (t11 t22 t33 t12 t13 t23) at a
// point with gradients g = (ux uy uz vx vy vz wx wy wz).
__device__ __forceinline__ bool accept(const float (&t)[NOUT], const float (&g)[9]) {
  const float t11 = t[0], t22 = t[1], t33 = t[2], t12 = t[3], t13 = t[4], t23 = t[5];
  const float det = t11 * (t22 * t33 - t23 * t23) - t12 * (t12 * t33 - t23 * t13) +
                    t13 * (t12 * t23 - t22 * t13);
  const bool real = t11 >= 0.f && t22 >= 0.f && t33 >= 0.f && t12 * t12 <= t11 * t22 &&
                    t13 * t13 <= t11 * t33 && t23 * t23 <= t22 * t33 && det >= 0.f;
  const float eps = -(t11 * g[0] + t22 * g[4] + t33 * g[8] + t12 * (g[1] + g[3]) +
                      t13 * (g[2] + g[6]) + t23 * (g[5] + g[7]));
  return real && eps >= 0.f;
}

// ---------------------------------------------------------------------------
// Training
// ---------------------------------------------------------------------------
struct TrainArgs {
  const Net* net;          // published forward tables (global)
  const bf16x8* bw;        // backward tables [BTN*64]
  const float* x;          // normalised inputs, [11][cap]
  const float* t;          // normalised targets, [6][cap]
  int64_t cap, n;          // stride and number of stored samples
  int64_t off, cnt;        // cnt > 0: the samples off..off+cnt-1 in order
  int64_t batch;           // cnt == 0: batch samples drawn with replacement
  uint64_t seed;
  float gscale;            // 2 / (samples * 6)
  float* partial;          // [gridDim.x][NPARAM + 1]: gradient sums, squared error
};

constexpr int TW = 2;                   // waves per training workgroup
constexpr int TROW = 40;                // padded row (BF16) of the transposes
struct TrainLds {
  float zs[TW][3 * 2 * 16 * 64];        // pre-activations of layers 1..3, one tile
  __bf16 tdz[TW][64 * TROW];            // [unit][sample] output-side gradients
  __bf16 th[TW][64 * TROW];             // [unit][sample] input-side activations
};

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Store a D-layout tile pair (BF16 of v, units 0..63) into a [unit][sample] transpose.
__device__ __forceinline__ void store_t(__bf16* tb, const f32x16 (&v)[2], int lane) {
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int i = 0; i < 16; ++i) tb[d_unit(mt, i, lane >> 5) * TROW + (lane & 31)] = (__bf16)v[mt][i];
}

// 8 consecutive samples 16s+8(l>>5).. of unit row r of a transpose (zero rows >= nrow).
__device__ __forceinline__ bf16x8 load_t(const __bf16* tb, int r, int nrow, int s, int lane) {
  if (r >= nrow) return bf16x8{};
  return *(const bf16x8*)(tb + r * TROW + 16 * s + 8 * (lane >> 5));
}

template <int NMT, int NNT>
__device__ __forceinline__ void grad_mfma(const __bf16* tdz, int nrow_o, const __bf16* th, int nrow_i,
                                          f32x16 (&acc)[NMT * NNT], int lane) {
#pragma unroll
  for (int s = 0; s < 2; ++s) {
#pragma unroll
    for (int mt = 0; mt < NMT; ++mt) {
      const bf16x8 a = load_t(tdz, 32 * mt + (lane & 31), nrow_o, s, lane);
#pragma unroll
      for (int nt = 0; nt < NNT; ++nt) {
        const bf16x8 b = load_t(th, 32 * nt + (lane & 31), nrow_i, s, lane);
        acc[mt * NNT + nt] = mfma(a, b, acc[mt * NNT + nt]);
      }
    }
  }
}

// Bias gradient: lane u < nrow adds the BF16 row u of the transpose.
__device__ __forceinline__ void bias_sum(const __bf16* tdz, int nrow, float& acc, int lane) {
  if (lane < nrow) {
#pragma unroll
    for (int p = 0; p < 32; ++p) acc += (float)tdz[lane * TROW + p];
  }
}

__global__ void __launch_bounds__(64 * TW) k_mlp_grad(TrainArgs a) {
  __shared__ TrainLds L;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, half = lane >> 5;
  float* zs = L.zs[w];
  __bf16* tdz = L.tdz[w];
  __bf16* th = L.th[w];
  const Net* net = a.net;
  f32x16 g1[2], g2[4], g3[4], g4[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) { g1[q] = f32x16{}; g4[q] = f32x16{}; }
#pragma unroll
  for (int q = 0; q < 4; ++q) { g2[q] = f32x16{}; g3[q] = f32x16{}; }
  float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, gb4 = 0.f, sq = 0.f;
  const int64_t total = a.cnt > 0 ? a.cnt : a.batch;
  const int64_t nchunk = (total + 63) / 64;
  for (int64_t c = (int64_t)blockIdx.x * TW + w; c < nchunk; c += (int64_t)gridDim.x * TW) {
    const int64_t pos = c * 64 + lane;
    const bool valid = pos < total;
    int64_t idx = 0;
    if (valid) idx = a.cnt > 0 ? a.off + pos : (int64_t)(mix64(a.seed ^ (uint64_t)pos) % (uint64_t)a.n);
    float x[16], tg[NOUT];
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = (valid && k < NIN) ? a.x[(int64_t)k * a.cap + idx] : 0.f;
#pragma unroll
    for (int k = 0; k < NOUT; ++k) tg[k] = valid ? a.t[(int64_t)k * a.cap + idx] : 0.f;
    bf16x8 bin[2];
    input_operands(x, lane, bin[0], bin[1]);
#pragma unroll 1
    for (int nt = 0; nt < 2; ++nt) {
      const f32x16 Y = forward_tile<true>(net, bin[nt], lane, zs);
      // targets and validity of the samples whose outputs this lane holds:
      // lanes < 32 hold units 0..3, lanes >= 32 units 4,5 of sample 32nt + (lane&31)
      float tt[4];
      const bool own = (nt == 0) == (half == 0);
      float sendt[4] = {half ? tg[0] : tg[4], half ? tg[1] : tg[5], half ? tg[2] : 0.f,
                        half ? tg[3] : 0.f};
      float recvt[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) recvt[q] = __shfl_xor(sendt[q], 32, 64);
      const int vs = __shfl_xor((int)valid, 32, 64);
      const bool vok = own ? valid : (vs != 0);
      if (!half) {
        for (int q = 0; q < 4; ++q) tt[q] = own ? tg[q] : recvt[q];
      } else {
        tt[0] = own ? tg[4] : recvt[0];
        tt[1] = own ? tg[5] : recvt[1];
        tt[2] = tt[3] = 0.f;
      }
      f32x16 dY = f32x16{};
      const int nu = half ? 2 : 4;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q < nu && vok) {
          const float r = Y[q] - tt[q];
          sq += r * r;
          dY[q] = a.gscale * r;
        }
      }
      // ---- layer 4: dW4, db4, dH3 = W4^T dY
      wave_fence();
      {
        f32x16 h3[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
          for (int i = 0; i < 16; ++i) h3[mt][i] = gelu(zs[((2 * 2 + mt) * 16 + i) * 64 + lane]);
        store_t(th, h3, lane);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (q < nu) tdz[(4 * half + q) * TROW + (lane & 31)] = (__bf16)dY[q];
      }
      wave_fence();
      grad_mfma<1, 2>(tdz, NOUT, th, 64, g4, lane);
      bias_sum(tdz, NOUT, gb4, lane);
      f32x16 dH[2];
      {
        const float r0 = __shfl_xor(dY[0], 32, 64), r1 = __shfl_xor(dY[1], 32, 64);
        bf16x8 bd = bf16x8{};
        if (!half) {
          bd[0] = (__bf16)dY[0]; bd[1] = (__bf16)dY[1]; bd[2] = (__bf16)dY[2];
          bd[3] = (__bf16)dY[3]; bd[4] = (__bf16)r0; bd[5] = (__bf16)r1;
        }
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) dH[mt] = mfma(a.bw[(BT4 + mt) * 64 + lane], bd, f32x16{});
      }
      // ---- layers 3, 2, 1
#pragma unroll
      for (int layer = 2; layer >= 0; --layer) {
        f32x16 dZ[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
          for (int i = 0; i < 16; ++i)
            dZ[mt][i] = (float)(__bf16)(dH[mt][i] * gelu_grad(zs[((layer * 2 + mt) * 16 + i) * 64 + lane]));
        wave_fence();
        store_t(tdz, dZ, lane);
        if (layer > 0) {
          f32x16 hin[2];
#pragma unroll
          for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int i = 0; i < 16; ++i) hin[mt][i] = gelu(zs[(((layer - 1) * 2 + mt) * 16 + i) * 64 + lane]);
          store_t(th, hin, lane);
        } else {
          // the layer-1 input: features 8half..8half+7 of sample lane&31
#pragma unroll
          for (int j = 0; j < 8; ++j) th[(8 * half + j) * TROW + (lane & 31)] = bin[nt][j];
        }
        wave_fence();
        if (layer == 2) {
          grad_mfma<2, 2>(tdz, 64, th, 64, g3, lane);
          bias_sum(tdz, 64, gb3, lane);
        } else if (layer == 1) {
          grad_mfma<2, 2>(tdz, 64, th, 64, g2, lane);
          bias_sum(tdz, 64, gb2, lane);
        } else {
          grad_mfma<2, 1>(tdz, 64, th, 16, g1, lane);
          bias_sum(tdz, 64, gb1, lane);
        }
        if (layer > 0) {
          bf16x8 B[4];
#pragma unroll
          for (int mt = 0; mt < 2; ++mt) {
            B[2 * mt] = pack8(dZ[mt], 0);
            B[2 * mt + 1] = pack8(dZ[mt], 8);
          }
          const int bt = layer == 2 ? BT3 : BT2;
#pragma unroll
          for (int mt = 0; mt < 2; ++mt) {
            dH[mt] = f32x16{};
#pragma unroll
            for (int s = 0; s < 4; ++s) dH[mt] = mfma(a.bw[(bt + mt * 4 + s) * 64 + lane], B[s], dH[mt]);
          }
        }
        wave_fence();
      }
    }
  }
  // ---- workgroup sum of the wave partials (in wave order), then out
  __syncthreads();
  float* red = &L.zs[0][0];
  static_assert(sizeof(L.zs) >= sizeof(float) * (NPARAM + 1), "reduction fits");
  for (int q = threadIdx.x; q < NPARAM + 1; q += 64 * TW) red[q] = 0.f;
  __syncthreads();
  for (int ww = 0; ww < TW; ++ww) {
    if (w == ww) {
      auto put = [&](const f32x16* acc, int nmt, int nnt, int ow, int mo, int ki) {
        for (int mt = 0; mt < nmt; ++mt)
          for (int nt = 0; nt < nnt; ++nt)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
              const int o = d_unit(mt, i, half), in = 32 * nt + (lane & 31);
              if (o < mo && in < ki) red[ow + o * ki + in] += acc[mt * nnt + nt][i];
            }
      };
      put(g1, 2, 1, OW1, NH, NIN);
      put(g2, 2, 2, OW2, NH, NH);
      put(g3, 2, 2, OW3, NH, NH);
      put(g4, 1, 2, OW4, NOUT, NH);
      red[OB1 + lane] += gb1;
      red[OB2 + lane] += gb2;
      red[OB3 + lane] += gb3;
      if (lane < NOUT) red[OB4 + lane] += gb4;
      float s = sq;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (lane == 0) red[NPARAM] += s;
    }
    __syncthreads();
  }
  float* out = a.partial + (size_t)blockIdx.x * (NPARAM + 1);
  for (int q = threadIdx.x; q < NPARAM + 1; q += 64 * TW) out[q] = red[q];
}

// Sum of the workgroup partials, then Adam on the FP32 master weights.
struct AdamArgs {
  const float* partial;
  int nblk;
  float* w;                // master weights [NPARAM]
  float* m;
  float* v;
  float* grad;             // out: summed gradient [NPARAM]
  float* loss;             // out: mean squared error of this step (may be null)
  float lr, b1, b2, eps, bc1, bc2;   // bc = 1 - beta^t
  float inv_n6;            // 1 / (samples * 6)
  int update;              // 0: gradient only
};

__global__ void __launch_bounds__(256) k_mlp_adam(AdamArgs a) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p > NPARAM) return;
  float g = 0.f;
  for (int b = 0; b < a.nblk; ++b) g += a.partial[(size_t)b * (NPARAM + 1) + p];
  if (p == NPARAM) {
    if (a.loss) *a.loss = g * a.inv_n6;
    return;
  }
  a.grad[p] = g;
  if (!a.update) return;
  const float m = a.b1 * a.m[p] + (1.f - a.b1) * g;
  const float v = a.b2 * a.v[p] + (1.f - a.b2) * g * g;
  a.m[p] = m;
  a.v[p] = v;
  a.w[p] -= a.lr * (m / a.bc1) / (sqrtf(v / a.bc2) + a.eps);
}

// Build the swizzled BF16 tables and the D-layout biases from the master weights.
__global__ void __launch_bounds__(256) k_mlp_publish(const float* w, Net* net, bf16x8* bw) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  constexpr int NF = FTN * 64 * 8, NB = BTN * 64 * 8;
  if (t < NF) {
    const int tile = t / 512, l = (t / 8) % 64, j = t % 8, hl = l >> 5;
    float val = 0.f;
    if (tile < FT2) {
      const int o = 32 * tile + (l & 31), in = 8 * hl + j;
      if (in < NIN) val = w[OW1 + o * NIN + in];
    } else if (tile < FT4) {
      const int tt = tile < FT3 ? tile - FT2 : tile - FT3, mt = tt / 4, s = tt % 4;
      const int o = 32 * mt + (l & 31), in = chain_unit(s, hl, j);
      val = w[(tile < FT3 ? OW2 : OW3) + o * NH + in];
    } else {
      const int s = tile - FT4, o = l & 31, in = chain_unit(s, hl, j);
      if (o < NOUT) val = w[OW4 + o * NH + in];
    }
    ((__bf16*)net->fw)[t] = (__bf16)val;
  } else if (t < NF + NB) {
    const int u = t - NF;
    const int tile = u / 512, l = (u / 8) % 64, j = u % 8, hl = l >> 5;
    float val = 0.f;
    if (tile < BT3) {
      const int m = 32 * tile + (l & 31), k = 8 * hl + j;
      if (k < NOUT) val = w[OW4 + k * NH + m];
    } else {
      const int tt = tile < BT2 ? tile - BT3 : tile - BT2, mt = tt / 4, s = tt % 4;
      const int m = 32 * mt + (l & 31), k = chain_unit(s, hl, j);
      val = w[(tile < BT2 ? OW3 : OW2) + k * NH + m];
    }
    ((__bf16*)bw)[u] = (__bf16)val;
  } else if (t < NF + NB + FBN) {
    const int u = t - NF - NB;
    float val = 0.f;
    if (u < FB4) {
      const int layer = u / 64, r = u % 64, mt = r / 32, hl = (r / 16) % 2, i = r % 16;
      val = w[(layer == 0 ? OB1 : layer == 1 ? OB2 : OB3) + d_unit(mt, i, hl)];
    } else {
      const int r = u - FB4, hl = r / 16, i = r % 16, o = d_unit(0, i, hl);
      if (o < NOUT) val = w[OB4 + o];
    }
    net->fb[u] = val;
  }
}

// Inference over n samples given as raw features [n][11]; one wave per 64.
__global__ void __launch_bounds__(64) k_mlp_forward(const Net* __restrict__ gnet, const float* f,
                                                    float* tau, int64_t n) {
  __shared__ Net net;
  const int lane = threadIdx.x;
  for (int q = lane; q < (int)(sizeof(Net) / 16); q += 64)
    ((uint4*)&net)[q] = ((const uint4*)gnet)[q];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  float x[NIN], y[NOUT];
#pragma unroll
  for (int k = 0; k < NIN; ++k) x[k] = i < n ? f[i * NIN + k] : 0.f;
  infer(&net, x, y, lane);
  if (i < n)
#pragma unroll
    for (int k = 0; k < NOUT; ++k) tau[i * NOUT + k] = y[k];
}

// ---------------------------------------------------------------------------
// Training data: box filter of a fine-grid state onto a coarse grid
// ---------------------------------------------------------------------------
// Input: physical u, v, e at levels 1..nz, w at faces 1..nz (w(nz) = 0), and
// the spectral derivatives ux uy vx vy (levels) and wx wy (faces), each
// [nz][ny][nx].  The coarse cell (I, J, K) covers fine points fx*I .. fx*I+fx-1,
// fy*J .., and levels fz*K+1 .. fz*K+fz.  Output [NFILT][ncz][ncy][ncx], all
// box means:
//   0 u, 1 v, 2 w (cell centre), 3 w at the coarse face K+1 (fine face
//   fz*(K+1), trapezoid over the fine faces fz*(K+1) -/+ fz/2, zero at the
//   top wall), 4 e (cell centre), 5..10 uu vv ww uv uw vw;
//   11..16 the fine grid's own SGS stresses t11 t22 t33 t12 t13 t23 at cell
//   centres.  These follow the TKE closure (complen.f90:20-21,
//   compsgs1.f90:41-50), with K = min(0.1 alk sqrt(e), xkmax) at faces:
//   -(K(k-1)+K(k)) ux, ..., and the face stresses -K (uz + wx) averaged to
//   the centre (the wall faces take the adjacent interior face's value).
// Cell-centred w is (w(k-1)+w(k))/2 with w(0) = 0; cell-centred e is
// (e(k-1)+e(k))/2 with e(0) = e(1) (lowerbc.f90:15).
constexpr int NFILT = 17;
struct FilterArgs {
  const double* u;
  const double* v;
  const double* w;
  const double* e;
  const double* ux;
  const double* uy;
  const double* vx;
  const double* vy;
  const double* wx;
  const double* wy;
  double* out;
  int nx, ny, nz, fx, fy, fz;
  double dz, alk, xkmax;
};

__global__ void __launch_bounds__(256) k_box_filter(FilterArgs a) {
  const int ncx = a.nx / a.fx, ncy = a.ny / a.fy, ncz = a.nz / a.fz;
  const int64_t nc = (int64_t)ncx * ncy * ncz;
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= nc) return;
  const int I = (int)(c % ncx), J = (int)((c / ncx) % ncy), K = (int)(c / ((int64_t)ncx * ncy));
  const size_t plane = (size_t)a.nx * a.ny;
  const double dzm = 1.0 / a.dz;
  auto at = [&](const double* f, int lev, size_t p) { return f[(size_t)(lev - 1) * plane + p]; };
  auto kf = [&](int face, size_t p) {          // eddy diffusivity at face 0..nz
    const double e = at(a.e, face < 1 ? 1 : face, p);
    const double k = 0.1 * a.alk * sqrt(e);
    return k < a.xkmax ? k : a.xkmax;
  };
  auto t13f = [&](int face, size_t p, double& t13, double& t23) {   // face 1..nz-1
    const double kk = kf(face, p);
    const double uz = (at(a.u, face + 1, p) - at(a.u, face, p)) * dzm;
    const double vz = (at(a.v, face + 1, p) - at(a.v, face, p)) * dzm;
    t13 = -kk * (uz + at(a.wx, face, p));
    t23 = -kk * (vz + at(a.wy, face, p));
  };
  double s[NFILT] = {};
  for (int kk = 0; kk < a.fz; ++kk) {
    const int k = a.fz * K + kk + 1;                    // fine level 1..nz
    for (int jj = 0; jj < a.fy; ++jj) {
      const size_t rowo = (size_t)(a.fy * J + jj) * a.nx;
      for (int ii = 0; ii < a.fx; ++ii) {
        const size_t p = rowo + a.fx * I + ii;
        const double uu = at(a.u, k, p);
        const double vv = at(a.v, k, p);
        const double wlo = k >= 2 ? at(a.w, k - 1, p) : 0.0;
        const double whi = at(a.w, k, p);
        const double wc = 0.5 * (wlo + whi);
        const double ec = 0.5 * (at(a.e, k >= 2 ? k - 1 : 1, p) + at(a.e, k, p));
        s[0] += uu; s[1] += vv; s[2] += wc; s[4] += ec;
        s[5] += uu * uu; s[6] += vv * vv; s[7] += wc * wc;
        s[8] += uu * vv; s[9] += uu * wc; s[10] += vv * wc;
        const double xs = kf(k - 1, p) + kf(k, p);
        s[11] += -xs * at(a.ux, k, p);
        s[12] += -xs * at(a.vy, k, p);
        s[13] += -xs * (whi - wlo) * dzm;
        s[14] += -xs * 0.5 * (at(a.uy, k, p) + at(a.vx, k, p));
        double a13, a23, b13, b23;
        t13f(k - 1 < 1 ? 1 : (k - 1 > a.nz - 1 ? a.nz - 1 : k - 1), p, a13, a23);
        t13f(k > a.nz - 1 ? a.nz - 1 : k, p, b13, b23);
        s[15] += 0.5 * (a13 + b13);
        s[16] += 0.5 * (a23 + b23);
      }
    }
  }
  // face filter: fine faces kf-h .. kf+h, end weights halved (h = fz/2)
  const int kfc = a.fz * (K + 1);
  double wf = 0.0, wsum = 0.0;
  if (K + 1 < ncz) {
    const int h = a.fz / 2;
    for (int d = -h; d <= h; ++d) {
      const double wt = (a.fz > 1 && (d == -h || d == h)) ? 0.5 : 1.0;
      const int k = kfc + d;                                // fine face, 1..nz-1 here
      for (int jj = 0; jj < a.fy; ++jj)
        for (int ii = 0; ii < a.fx; ++ii)
          wf += wt * at(a.w, k, (size_t)(a.fy * J + jj) * a.nx + a.fx * I + ii);
      wsum += wt;
    }
    wf /= wsum * a.fx * a.fy;
  }
  const double inv = 1.0 / ((double)a.fx * a.fy * a.fz);
  for (int q = 0; q < NFILT; ++q) a.out[(size_t)q * nc + c] = q == 3 ? wf : s[q] * inv;
}

}  // namespace mlp
}  // namespace cf
