// Warm-start MLP 12 -> 64 -> 64 -> 2 on gfx950 BF16 MFMA (no reference
// counterpart: the reference starts every substep's multiplier at zero,
// simkernels.js:146).
//
// The model (every implementation, the oracle included, computes exactly this):
//   x = bf16(features[0..11]), padded with zeros to 16
//   h1 = bf16(relu(W1 x + b1))          W1 64x12, f32 accumulation
//   h2 = bf16(relu(W2 h1 + b2))         W2 64x64
//   y  = W3 h2 + b3                     W3 2x64, f32
// bf16 values are round-to-nearest-even of f32; weights are bf16 copies of f32
// master weights.  Canonical parameter order (the C ABI's float vector):
//   W1[64][12], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]   (row = output unit)
//
// Layers run in transposed form, one wave per 32 samples, on
// v_mfma_f32_32x32x16_bf16: for A (32xK), B (Kx32) and D (32x32), lane l holds
// A[l%32][8(l/32) + j], B[8(l/32) + j][l%32] (j < 8) and D register r holds
// D[8(r/4) + 4(l/32) + r%4][l%32].  With samples as the columns, a layer's D
// registers are the next layer's B operand once the K index is permuted: the
// registers 8s..8s+7 of output tile t hold units 32t + 16s + 8(j/4) + 4(l/32) + j%4,
// and the weights are stored pre-permuted to match, so activations never leave
// registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xpbi {

constexpr int kMlpIn = 12;
constexpr int kMlpH = 64;
constexpr int kMlpOut = 2;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;   // 768
constexpr int kOffW2 = kOffB1 + kMlpH;            // 832
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;    // 4928
constexpr int kOffW3 = kOffB2 + kMlpH;            // 4992
constexpr int kOffB3 = kOffW3 + kMlpOut * kMlpH;  // 5120
constexpr int kMlpParams = kOffB3 + kMlpOut;      // 5122
constexpr int kFeatStride = 16;                   // floats per feature row (12 features + 4 extra)

// bf16 round-to-nearest-even (NaN kept quiet)
__host__ __device__ inline uint16_t bf16_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_float(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
__host__ __device__ inline float bf16_round(float f) { return bf16_float(bf16_bits(f)); }

// packed bf16 weights of one network (what a kernel reads): the forward part,
// which inference stages alone, then the transposed layer 2 of the backward pass
struct alignas(16) MlpFwd {
  uint16_t w1[kMlpH * 16];     // [unit][feature], features 12..15 zero
  uint16_t w2f[kMlpH * kMlpH];  // forward A operand of layer 2, [t'][t][s][lane][j]
  float w3[kMlpOut * kMlpH];    // bf16-rounded, [k][unit]
  float b1[kMlpH], b2[kMlpH], b3[4];
};
struct alignas(16) MlpPacked {
  MlpFwd f;
  uint16_t w2b[kMlpH * kMlpH];  // A operand of dh1 = W2^T dz2, [ti][t'][s][lane][j]
};

// unit of register j (< 8) of lane half h in K-step s of 32-unit tile t
__host__ __device__ inline int mlp_kunit(int t, int s, int h, int j) { return 32 * t + 16 * s + 8 * (j / 4) + 4 * h + j % 4; }

// slot of W2[o][i] in w2f and w2b
__host__ __device__ inline void mlp_w2_slots(int o, int i, int* fwd, int* bwd) {
  // forward: row o (output tile o/32, lane o%32 + 32h), K unit i
  {
    const int t = i / 32, r = i % 32, s = r / 16, r2 = r % 16, hi = r2 / 8, rem = r2 % 8;
    const int h = rem / 4, j = 4 * hi + rem % 4;
    const int lane = o % 32 + 32 * h;
    *fwd = ((((o / 32) * 2 + t) * 2 + s) * 64 + lane) * 8 + j;
  }
  // backward: row i (tile i/32, lane i%32 + 32h), K unit o
  {
    const int t = o / 32, r = o % 32, s = r / 16, r2 = r % 16, hi = r2 / 8, rem = r2 % 8;
    const int h = rem / 4, j = 4 * hi + rem % 4;
    const int lane = i % 32 + 32 * h;
    *bwd = ((((i / 32) * 2 + t) * 2 + s) * 64 + lane) * 8 + j;
  }
}

// write canonical parameter k (value v) into the packed copy
__host__ __device__ inline void mlp_pack_one(MlpPacked* P, int k, float v) {
  if (k < kOffB1) {
    const int u = k / kMlpIn, f = k % kMlpIn;
    P->f.w1[u * 16 + f] = bf16_bits(v);
  } else if (k < kOffW2) {
    P->f.b1[k - kOffB1] = v;
  } else if (k < kOffB2) {
    const int q = k - kOffW2, o = q / kMlpH, i = q % kMlpH;
    int a, b;
    mlp_w2_slots(o, i, &a, &b);
    P->f.w2f[a] = bf16_bits(v);
    P->w2b[b] = bf16_bits(v);
  } else if (k < kOffW3) {
    P->f.b2[k - kOffB2] = v;
  } else if (k < kOffB3) {
    P->f.w3[k - kOffW3] = bf16_round(v);
  } else {
    P->f.b3[k - kOffB3] = v;
  }
}

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ bf16x8 ld_bf16x8(const uint16_t* p) {
  bf16x8 v;
  __builtin_memcpy(&v, p, 16);
  return v;
}
__device__ __forceinline__ __bf16 to_bf16(float f) {
  const uint16_t b = bf16_bits(f);
  __bf16 r;
  __builtin_memcpy(&r, &b, 2);
  return r;
}
__device__ __forceinline__ f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
// the unit held in register r of lane half h for 32-unit tile t (D layout)
__device__ __forceinline__ int mlp_dunit(int t, int r, int h) { return 32 * t + 8 * (r / 4) + 4 * h + (r % 4); }

// activations of one wave's 32 samples, sample l%32 in lane l
struct MlpAct {
  bf16x8 xb;         // features 8h..8h+7
  bf16x8 h1[2][2];   // [tile][K-step]
  float h2[2][16];   // bf16-rounded, D layout
  uint32_t pos1, pos2;  // relu masks, bit 16t + r
  float y[2];
};

// forward of the wave's 32 samples; xb = this lane's 8 features (8h..8h+7, zero past 11)
__device__ __forceinline__ void mlp_forward_b(const MlpFwd& W, bf16x8 xb, MlpAct& a) {
  const int l = threadIdx.x & 63, h = l >> 5;
  a.xb = xb;
  a.pos1 = 0u;
  a.pos2 = 0u;
#pragma unroll
  for (int t = 0; t < 2; t++) {
    f32x16 d = {};
    d = mfma32(ld_bf16x8(&W.w1[(32 * t + (l & 31)) * 16 + 8 * h]), a.xb, d);
#pragma unroll
    for (int s = 0; s < 2; s++) {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int r = 8 * s + j;
        const float z = d[r] + W.b1[mlp_dunit(t, r, h)];
        if (z > 0.0f) a.pos1 |= 1u << (16 * t + r);
        a.h1[t][s][j] = to_bf16(z > 0.0f ? z : 0.0f);
      }
    }
  }
#pragma unroll
  for (int to = 0; to < 2; to++) {
    f32x16 d = {};
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
      for (int s = 0; s < 2; s++) d = mfma32(ld_bf16x8(&W.w2f[(((to * 2 + t) * 2 + s) * 64 + l) * 8]), a.h1[t][s], d);
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const float z = d[r] + W.b2[mlp_dunit(to, r, h)];
      if (z > 0.0f) a.pos2 |= 1u << (16 * to + r);
      a.h2[to][r] = bf16_round(z > 0.0f ? z : 0.0f);
    }
  }
#pragma unroll
  for (int k = 0; k < 2; k++) {
    float acc = 0.0f;
#pragma unroll
    for (int to = 0; to < 2; to++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc += W.w3[k * kMlpH + mlp_dunit(to, r, h)] * a.h2[to][r];
    acc += __shfl_xor(acc, 32, 64);
    a.y[k] = acc + W.b3[k];
  }
}
// the same with the features as f32
__device__ __forceinline__ void mlp_forward(const MlpFwd& W, const float x8[8], MlpAct& a) {
  bf16x8 xb;
#pragma unroll
  for (int j = 0; j < 8; j++) xb[j] = to_bf16(x8[j]);
  mlp_forward_b(W, xb, a);
}
// stage the forward weights into LDS (the caller synchronises the block)
__device__ __forceinline__ void mlp_stage_fwd(MlpFwd& s, const MlpFwd* __restrict__ g) {
  const uint4* src = (const uint4*)g;
  uint4* dst = (uint4*)&s;
  for (uint32_t k = threadIdx.x; k < sizeof(MlpFwd) / 16; k += blockDim.x) dst[k] = src[k];
}
#endif

}  // namespace xpbi
