// Warm-start MLP: inference, training (forward + backward on BF16 MFMA with f32
// accumulation, deterministic two-stage gradient reduction), Adam, and the
// trainer's C ABI.  The model and the register layouts are defined in
// xpbi_mlp.hpp.
//
// Backward of the loss L = (1/B) sum_n w_n sum_k m_nk (y_nk - t_nk)^2 (w: the
// sample weights, 1 without; m = 0 where the target is NaN), with every operand
// of a weight-gradient product rounded to bf16 and f32 accumulation:
//   dy = (2/B) w m (y - t);  dW3 = sum bf16(dy) h2^T;  db3 = sum bf16(dy)
//   dz2 = (W3^T dy) [z2 > 0];  dW2 = sum bf16(dz2) h1^T;  db2 = sum bf16(dz2)
//   dz1 = (W2^T bf16(dz2)) [z1 > 0];  dW1 = sum bf16(dz1) x^T;  db1 = sum bf16(dz1)
// A block of 4 waves takes 32 samples per wave per chunk; the per-sample
// factors are transposed through LDS so that the sums over samples run as the
// K dimension of MFMAs.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "xpbi_internal.hpp"
#include "xpbi_mlp.hpp"

namespace xpbi {

constexpr int kMlpWaves = 4;
constexpr int kMlpBlock = 64 * kMlpWaves;
constexpr int kGradLen = kMlpParams + 1;  // + loss

__device__ __forceinline__ void stage_weights(MlpPacked& s, const MlpPacked* __restrict__ g) {
  const uint4* src = (const uint4*)g;
  uint4* dst = (uint4*)&s;
  for (u32 k = threadIdx.x; k < sizeof(MlpPacked) / 16; k += blockDim.x) dst[k] = src[k];
  __syncthreads();
}

// this lane's 8 features (8h..8h+7) of row `row` of a [rows][16] array; zero past feature 11
__device__ __forceinline__ void load_x8(const float* __restrict__ X, u64 row, bool valid, float x8[8]) {
  const int h = (threadIdx.x & 63) >> 5;
  if (!valid) {
#pragma unroll
    for (int j = 0; j < 8; j++) x8[j] = 0.0f;
    return;
  }
  const float4* p = (const float4*)(X + row * kFeatStride + 8 * h);
  const float4 a = p[0], b = p[1];
  x8[0] = a.x; x8[1] = a.y; x8[2] = a.z; x8[3] = a.w;
  x8[4] = h ? 0.0f : b.x; x8[5] = h ? 0.0f : b.y; x8[6] = h ? 0.0f : b.z; x8[7] = h ? 0.0f : b.w;
}

// ---------------------------------------------------------------- inference
__global__ __launch_bounds__(kMlpBlock) void k_mlp_predict(const MlpFwd* __restrict__ Wg, const float* __restrict__ X,
                                                           u64 rows, float* __restrict__ Y) {
  __shared__ MlpFwd W;
  mlp_stage_fwd(W, Wg);
  __syncthreads();
  const u32 l = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const u64 base = ((u64)blockIdx.x * kMlpWaves + w) * 32u;
  const u64 n = base + (l & 31u);
  float x8[8];
  load_x8(X, n, n < rows, x8);
  MlpAct a;
  mlp_forward(W, x8, a);
  if (l < 32u && n < rows) {
    Y[2 * n] = a.y[0];
    Y[2 * n + 1] = a.y[1];
  }
}

// ---------------------------------------------------------------- training
struct alignas(16) WaveT {
  uint16_t h1[kMlpH * 32];   // [unit][sample]
  uint16_t h2[kMlpH * 32];
  uint16_t dz1[kMlpH * 32];
  uint16_t dz2[kMlpH * 32];
  uint16_t x[16 * 32];       // [feature][sample]
  float dy[kMlpOut * 32];    // bf16-rounded dy, [k][sample]
};

__global__ __launch_bounds__(kMlpBlock) void k_mlp_train(const MlpPacked* __restrict__ Wg, const float* __restrict__ X,
                                                         const float* __restrict__ T, const float* __restrict__ Wt,
                                                         const u32* __restrict__ idx, u64 first, u32 count, u32 chunks,
                                                         float* __restrict__ part) {
  __shared__ MlpPacked W;
  __shared__ WaveT S[kMlpWaves];
  __shared__ float acc[kGradLen];
  stage_weights(W, Wg);
  for (u32 k = threadIdx.x; k < kGradLen; k += blockDim.x) acc[k] = 0.0f;
  const u32 l = threadIdx.x & 63u, w = threadIdx.x >> 6, n32 = l & 31u;
  const int h = (int)(l >> 5);
  WaveT& st = S[w];
  const float scale = 2.0f / (float)count;
  f32x16 g2[2][2] = {}, g1[2] = {};
  float gb1 = 0.0f, gb2 = 0.0f, gw3[2] = {0.0f, 0.0f}, gb3 = 0.0f, loss = 0.0f;
  for (u32 c = 0; c < chunks; c++) {
    const u32 q = ((blockIdx.x * chunks + c) * kMlpWaves + w) * 32u + n32;  // sample within the batch
    const bool valid = q < count;
    const u64 row = valid ? (idx ? (u64)idx[first + q] : first + q) : 0;
    float x8[8];
    load_x8(X, row, valid, x8);
    MlpAct a;
    mlp_forward(W.f, x8, a);
    // output error
    float dy[2];
    const float wn = (valid && Wt) ? Wt[row] : 1.0f;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const float t = valid ? T[2 * row + k] : NAN;
      const bool m = valid && !isnan(t);
      const float e = m ? a.y[k] - t : 0.0f;
      if (h == 0) loss += wn * e * e;
      dy[k] = scale * wn * e;
    }
    // dz2 = (W3^T dy) [z2 > 0], D layout
    bf16x8 dz2b[2][2];
#pragma unroll
    for (int to = 0; to < 2; to++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int u = mlp_dunit(to, r, h);
        float v = W.f.w3[u] * dy[0] + W.f.w3[kMlpH + u] * dy[1];
        if (!((a.pos2 >> (16 * to + r)) & 1u)) v = 0.0f;
        dz2b[to][r / 8][r % 8] = to_bf16(v);
        st.dz2[u * 32 + n32] = bf16_bits(v);
        st.h2[u * 32 + n32] = bf16_bits(a.h2[to][r]);
      }
    // dz1 = (W2^T bf16(dz2)) [z1 > 0]
#pragma unroll
    for (int ti = 0; ti < 2; ti++) {
      f32x16 d = {};
#pragma unroll
      for (int to = 0; to < 2; to++)
#pragma unroll
        for (int s = 0; s < 2; s++) d = mfma32(ld_bf16x8(&W.w2b[(((ti * 2 + to) * 2 + s) * 64 + l) * 8]), dz2b[to][s], d);
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int u = mlp_dunit(ti, r, h);
        const float v = ((a.pos1 >> (16 * ti + r)) & 1u) ? d[r] : 0.0f;
        st.dz1[u * 32 + n32] = bf16_bits(v);
      }
    }
    // h1 and x, transposed
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
      for (int s = 0; s < 2; s++) {
        uint16_t b[8];
        __builtin_memcpy(b, &a.h1[t][s], 16);
#pragma unroll
        for (int j = 0; j < 8; j++) st.h1[mlp_kunit(t, s, h, j) * 32 + n32] = b[j];
      }
    {
      uint16_t b[8];
      __builtin_memcpy(b, &a.xb, 16);
#pragma unroll
      for (int j = 0; j < 8; j++) st.x[(8 * h + j) * 32 + n32] = b[j];
    }
    if (h == 0) {
      st.dy[n32] = bf16_round(dy[0]);
      st.dy[32 + n32] = bf16_round(dy[1]);
    }
    __syncthreads();
    // weight gradients: sums over the wave's 32 samples as the MFMA K dimension
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const int ko = 16 * s + 8 * h;
#pragma unroll
      for (int to = 0; to < 2; to++) {
        const bf16x8 A = ld_bf16x8(&st.dz2[(32 * to + n32) * 32 + ko]);
#pragma unroll
        for (int ti = 0; ti < 2; ti++) g2[to][ti] = mfma32(A, ld_bf16x8(&st.h1[(32 * ti + n32) * 32 + ko]), g2[to][ti]);
      }
      bf16x8 B;
      if (n32 < (u32)kMlpIn) B = ld_bf16x8(&st.x[n32 * 32 + ko]);
      else {
#pragma unroll
        for (int j = 0; j < 8; j++) B[j] = to_bf16(0.0f);
      }
#pragma unroll
      for (int tu = 0; tu < 2; tu++) g1[tu] = mfma32(ld_bf16x8(&st.dz1[(32 * tu + n32) * 32 + ko]), B, g1[tu]);
    }
    // bias gradients and dW3 (lane l = unit l)
    {
      float s1 = 0.0f, s2 = 0.0f, w0 = 0.0f, w1 = 0.0f;
      for (int n = 0; n < 32; n++) {
        const float hv = bf16_float(st.h2[l * 32 + n]);
        s1 += bf16_float(st.dz1[l * 32 + n]);
        s2 += bf16_float(st.dz2[l * 32 + n]);
        w0 += st.dy[n] * hv;
        w1 += st.dy[32 + n] * hv;
      }
      gb1 += s1;
      gb2 += s2;
      gw3[0] += w0;
      gw3[1] += w1;
      if (l < 2u) {
        float s3 = 0.0f;
        for (int n = 0; n < 32; n++) s3 += st.dy[32 * l + n];
        gb3 += s3;
      }
    }
    __syncthreads();
  }
  // loss over the wave
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
  // block sum in a fixed wave order
  for (u32 ww = 0; ww < (u32)kMlpWaves; ww++) {
    if (w == ww) {
#pragma unroll
      for (int to = 0; to < 2; to++)
#pragma unroll
        for (int ti = 0; ti < 2; ti++)
#pragma unroll
          for (int r = 0; r < 16; r++)
            acc[kOffW2 + mlp_dunit(to, r, h) * kMlpH + 32 * ti + n32] += g2[to][ti][r];
      if (n32 < (u32)kMlpIn) {
#pragma unroll
        for (int tu = 0; tu < 2; tu++)
#pragma unroll
          for (int r = 0; r < 16; r++) acc[kOffW1 + mlp_dunit(tu, r, h) * kMlpIn + n32] += g1[tu][r];
      }
      acc[kOffB1 + l] += gb1;
      acc[kOffB2 + l] += gb2;
      acc[kOffW3 + l] += gw3[0];
      acc[kOffW3 + kMlpH + l] += gw3[1];
      if (l < 2u) acc[kOffB3 + l] += gb3;
      if (l == 0u) acc[kMlpParams] += loss / (float)count;
    }
    __syncthreads();
  }
  float* out = part + (u64)blockIdx.x * kGradLen;
  for (u32 k = threadIdx.x; k < (u32)kGradLen; k += blockDim.x) out[k] = acc[k];
}

// grad[k] = sum over blocks (fixed order); the loss goes to lossLog[slot]
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ part, u32 nblk, float* __restrict__ grad,
                                                    float* __restrict__ lossLog, u32 slot) {
  const u32 k = blockIdx.x * 256u + threadIdx.x;
  if (k >= (u32)kGradLen) return;
  float s = 0.0f;
  for (u32 b = 0; b < nblk; b++) s += part[(u64)b * kGradLen + k];
  grad[k] = s;
  if (k == (u32)kMlpParams && lossLog) lossLog[slot] = s;
}

// Adam (Kingma & Ba 2015) on the f32 masters, then the packed copies
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ g, float lr, float b1, float b2, float eps,
                                                  float bc1, float bc2, MlpPacked* __restrict__ P) {
  const u32 k = blockIdx.x * 256u + threadIdx.x;
  if (k >= (u32)kMlpParams) return;
  const float gk = g[k];
  const float mk = b1 * m[k] + (1.0f - b1) * gk;
  const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  const float wk = w[k] - lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
  w[k] = wk;
  mlp_pack_one(P, (int)k, wk);
}

__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ w, MlpPacked* __restrict__ P) {
  const u32 k = blockIdx.x * 256u + threadIdx.x;
  if (k < (u32)kMlpParams) mlp_pack_one(P, (int)k, w[k]);
}

static inline u32 cdivu(u64 a, u64 b) { return (u32)((a + b - 1) / b); }

// host-side packing of canonical parameters (the sim's copy)
void mlp_pack_host(const float* params, void* packed) {
  MlpPacked* P = (MlpPacked*)packed;
  memset(P, 0, sizeof(MlpPacked));
  for (int k = 0; k < kMlpParams; k++) mlp_pack_one(P, k, params[k]);
}
size_t mlp_packed_bytes() { return sizeof(MlpPacked); }

}  // namespace xpbi

// ---------------------------------------------------------------- trainer C ABI
using namespace xpbi;

struct xpbi_mlp {
  int dev = 0;
  hipStream_t stream = nullptr;
  float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *part = nullptr, *lossLog = nullptr;
  MlpPacked* P = nullptr;
  float *X = nullptr, *T = nullptr, *Wt = nullptr;
  u32* idx = nullptr;
  size_t rows = 0, partBlocks = 0, lossCap = 0;
  long step = 0;
  std::string err;
};

namespace {
int mfail(xpbi_mlp* m, int code, const std::string& msg) {
  if (m) m->err = msg;
  return code;
}
int mhip(xpbi_mlp* m, hipError_t e, const char* what) {
  return mfail(m, XPBI_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
void mfree(void*& p) {
  if (p) hipFree(p);
  p = nullptr;
}
template <class T> hipError_t realloc_dev(T*& p, size_t count) {
  if (p) hipFree(p);
  p = nullptr;
  return hipMalloc((void**)&p, (count ? count : 1) * sizeof(T));
}
// chunks per wave for a batch of `count` samples: up to 8, with at least 64 blocks when possible
u32 chunks_for(u32 count, u32* blocks) {
  const u32 perChunk = 32u * kMlpWaves;
  u32 ch = 1;
  while (ch < 8 && cdivu(count, perChunk * ch * 2) >= 64) ch *= 2;
  *blocks = cdivu(count, perChunk * ch);
  return ch;
}
int ensure_part(xpbi_mlp* m, size_t blocks) {
  if (blocks <= m->partBlocks) return XPBI_OK;
  hipError_t e = realloc_dev(m->part, blocks * kGradLen);
  if (e != hipSuccess) return mhip(m, e, "hipMalloc");
  m->partBlocks = blocks;
  return XPBI_OK;
}
int ensure_loss(xpbi_mlp* m, size_t n) {
  if (n <= m->lossCap) return XPBI_OK;
  hipError_t e = realloc_dev(m->lossLog, n);
  if (e != hipSuccess) return mhip(m, e, "hipMalloc");
  m->lossCap = n;
  return XPBI_OK;
}
// loss and gradient of data rows idx[first .. first+count) (or first.. without idx) into m->grad
int grad_pass(xpbi_mlp* m, const u32* idx, size_t first, u32 count, u32 lossSlot) {
  u32 blocks;
  const u32 ch = chunks_for(count, &blocks);
  int r = ensure_part(m, blocks);
  if (r) return r;
  hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(kMlpBlock), 0, m->stream, (const MlpPacked*)m->P,
                     (const float*)m->X, (const float*)m->T, (const float*)m->Wt, idx, (u64)first, count, ch,
                     m->part);
  hipLaunchKernelGGL(k_mlp_reduce, dim3(cdivu(kGradLen, 256)), dim3(256), 0, m->stream, (const float*)m->part,
                     blocks, m->grad, m->lossLog, lossSlot);
  return XPBI_OK;
}
void adam(xpbi_mlp* m, float lr) {
  m->step++;
  const double b1 = 0.9, b2 = 0.999;
  const float bc1 = (float)(1.0 - pow(b1, (double)m->step)), bc2 = (float)(1.0 - pow(b2, (double)m->step));
  hipLaunchKernelGGL(k_mlp_adam, dim3(cdivu(kMlpParams, 256)), dim3(256), 0, m->stream, m->w, m->m, m->v,
                     (const float*)m->grad, lr, (float)b1, (float)b2, 1e-8f, bc1, bc2, m->P);
}
uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
int msync(xpbi_mlp* m) {
  hipError_t e = hipStreamSynchronize(m->stream);
  return e == hipSuccess ? XPBI_OK : mhip(m, e, "hipStreamSynchronize");
}
// host [rows][12] -> device [rows][16]
int upload_rows(xpbi_mlp* m, float** dst, const float* x, size_t rows) {
  std::vector<float> buf(rows * kFeatStride, 0.0f);
  for (size_t i = 0; i < rows; i++) memcpy(&buf[i * kFeatStride], x + i * kMlpIn, kMlpIn * sizeof(float));
  hipError_t e = realloc_dev(*dst, buf.size());
  if (e != hipSuccess) return mhip(m, e, "hipMalloc");
  e = hipMemcpy(*dst, buf.data(), buf.size() * sizeof(float), hipMemcpyHostToDevice);
  return e == hipSuccess ? XPBI_OK : mhip(m, e, "hipMemcpy");
}
}  // namespace

extern "C" {

int xpbi_mlp_create(int device, xpbi_mlp** out) {
  if (!out) return XPBI_ERR_ARG;
  *out = nullptr;
  xpbi_mlp* m = new xpbi_mlp();
  m->dev = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = realloc_dev(m->w, kMlpParams);
  if (e == hipSuccess) e = realloc_dev(m->m, kMlpParams);
  if (e == hipSuccess) e = realloc_dev(m->v, kMlpParams);
  if (e == hipSuccess) e = realloc_dev(m->grad, kGradLen);
  if (e == hipSuccess) e = realloc_dev(m->P, 1);
  if (e == hipSuccess) e = hipMemset(m->w, 0, kMlpParams * sizeof(float));
  if (e == hipSuccess) e = hipMemset(m->m, 0, kMlpParams * sizeof(float));
  if (e == hipSuccess) e = hipMemset(m->v, 0, kMlpParams * sizeof(float));
  if (e == hipSuccess) e = hipMemset(m->P, 0, sizeof(MlpPacked));
  if (e != hipSuccess) {
    xpbi_mlp_destroy(m);
    return XPBI_ERR_HIP;
  }
  *out = m;
  return XPBI_OK;
}

void xpbi_mlp_destroy(xpbi_mlp* m) {
  if (!m) return;
  hipSetDevice(m->dev);
  if (m->stream) hipStreamSynchronize(m->stream);
  void* ps[] = {m->w, m->m, m->v, m->grad, m->part, m->lossLog, m->P, m->X, m->T, m->Wt, m->idx};
  for (void* p : ps) mfree(p);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
}

const char* xpbi_mlp_last_error(const xpbi_mlp* m) { return m ? m->err.c_str() : "null mlp"; }

int xpbi_mlp_set_params(xpbi_mlp* m, const float* params, size_t n) {
  if (!m) return XPBI_ERR_ARG;
  if (!params || n != (size_t)kMlpParams) return mfail(m, XPBI_ERR_ARG, "params must hold XPBI_MLP_PARAMS floats");
  hipSetDevice(m->dev);
  int r = msync(m);
  if (r) return r;
  hipError_t e = hipMemcpy(m->w, params, n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(m->m, 0, n * sizeof(float));
  if (e == hipSuccess) e = hipMemset(m->v, 0, n * sizeof(float));
  if (e != hipSuccess) return mhip(m, e, "set_params");
  m->step = 0;
  hipLaunchKernelGGL(k_mlp_pack, dim3(cdivu(kMlpParams, 256)), dim3(256), 0, m->stream, (const float*)m->w, m->P);
  return msync(m);
}

int xpbi_mlp_get_params(xpbi_mlp* m, float* params, size_t n) {
  if (!m) return XPBI_ERR_ARG;
  if (!params || n != (size_t)kMlpParams) return mfail(m, XPBI_ERR_ARG, "params must hold XPBI_MLP_PARAMS floats");
  hipSetDevice(m->dev);
  int r = msync(m);
  if (r) return r;
  hipError_t e = hipMemcpy(params, m->w, n * sizeof(float), hipMemcpyDeviceToHost);
  return e == hipSuccess ? XPBI_OK : mhip(m, e, "get_params");
}

int xpbi_mlp_set_data(xpbi_mlp* m, const float* x, const float* t, const float* w, size_t rows) {
  if (!m) return XPBI_ERR_ARG;
  if (rows > 0 && (!x || !t)) return mfail(m, XPBI_ERR_ARG, "x and t must not be null");
  if (rows >= (1ull << 31)) return mfail(m, XPBI_ERR_ARG, "too many rows");
  hipSetDevice(m->dev);
  int r = msync(m);
  if (r) return r;
  if ((r = upload_rows(m, &m->X, x, rows))) return r;
  hipError_t e = realloc_dev(m->T, rows * 2);
  if (e == hipSuccess) e = hipMemcpy(m->T, t, rows * 2 * sizeof(float), hipMemcpyHostToDevice);
  if (m->Wt) hipFree(m->Wt);
  m->Wt = nullptr;
  if (e == hipSuccess && w) {
    e = realloc_dev(m->Wt, rows);
    if (e == hipSuccess) e = hipMemcpy(m->Wt, w, rows * sizeof(float), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = realloc_dev(m->idx, rows);
  if (e != hipSuccess) return mhip(m, e, "set_data");
  m->rows = rows;
  return XPBI_OK;
}

int xpbi_mlp_predict(xpbi_mlp* m, const float* x, size_t rows, float* y) {
  if (!m) return XPBI_ERR_ARG;
  if (rows == 0) return XPBI_OK;
  if (!x || !y) return mfail(m, XPBI_ERR_ARG, "x and y must not be null");
  hipSetDevice(m->dev);
  float* dX = nullptr;
  float* dY = nullptr;
  int r = upload_rows(m, &dX, x, rows);
  if (r) return r;
  hipError_t e = hipMalloc((void**)&dY, rows * 2 * sizeof(float));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mlp_predict, dim3(cdivu(rows, 32 * kMlpWaves)), dim3(kMlpBlock), 0, m->stream,
                       (const MlpFwd*)m->P, (const float*)dX, (u64)rows, dY);
    e = hipStreamSynchronize(m->stream);
  }
  if (e == hipSuccess) e = hipMemcpy(y, dY, rows * 2 * sizeof(float), hipMemcpyDeviceToHost);
  hipFree(dX);
  if (dY) hipFree(dY);
  return e == hipSuccess ? XPBI_OK : mhip(m, e, "predict");
}

int xpbi_mlp_loss_grad(xpbi_mlp* m, size_t first, size_t count, float* grad, float* loss) {
  if (!m) return XPBI_ERR_ARG;
  if (count == 0 || first + count > m->rows) return mfail(m, XPBI_ERR_ARG, "rows out of range of the data");
  hipSetDevice(m->dev);
  int r = ensure_loss(m, 1);
  if (r || (r = grad_pass(m, nullptr, first, (u32)count, 0))) return r;
  if ((r = msync(m))) return r;
  std::vector<float> g(kGradLen);
  hipError_t e = hipMemcpy(g.data(), m->grad, kGradLen * sizeof(float), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return mhip(m, e, "loss_grad");
  if (grad) memcpy(grad, g.data(), kMlpParams * sizeof(float));
  if (loss) *loss = g[kMlpParams];
  return XPBI_OK;
}

int xpbi_mlp_adam_step(xpbi_mlp* m, const float* grad, float lr) {
  if (!m) return XPBI_ERR_ARG;
  hipSetDevice(m->dev);
  if (grad) {
    hipError_t e = hipMemcpy(m->grad, grad, kMlpParams * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return mhip(m, e, "adam_step");
  }
  adam(m, lr);
  return msync(m);
}

int xpbi_mlp_train(xpbi_mlp* m, int epochs, int batch, float lr, uint64_t seed, float* epoch_loss) {
  if (!m) return XPBI_ERR_ARG;
  if (epochs < 0 || batch <= 0) return mfail(m, XPBI_ERR_ARG, "epochs >= 0 and batch > 0");
  if (m->rows == 0) return mfail(m, XPBI_ERR_ARG, "no training data");
  hipSetDevice(m->dev);
  const size_t steps = (m->rows + (size_t)batch - 1) / (size_t)batch;
  int r = ensure_loss(m, steps);
  if (r) return r;
  std::vector<u32> perm(m->rows);
  std::vector<float> losses(steps);
  uint64_t st = seed;
  for (int ep = 0; ep < epochs; ep++) {
    for (size_t i = 0; i < m->rows; i++) perm[i] = (u32)i;
    for (size_t i = m->rows - 1; i > 0; i--) std::swap(perm[i], perm[splitmix(st) % (i + 1)]);
    if ((r = msync(m))) return r;
    hipError_t e = hipMemcpy(m->idx, perm.data(), m->rows * sizeof(u32), hipMemcpyHostToDevice);
    if (e != hipSuccess) return mhip(m, e, "train");
    for (size_t k = 0; k < steps; k++) {
      const size_t first = k * (size_t)batch;
      const u32 cnt = (u32)std::min((size_t)batch, m->rows - first);
      if ((r = grad_pass(m, m->idx, first, cnt, (u32)k))) return r;
      adam(m, lr);
    }
    if ((r = msync(m))) return r;
    e = hipMemcpy(losses.data(), m->lossLog, steps * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return mhip(m, e, "train");
    if (epoch_loss) {
      double s = 0.0, wsum = 0.0;
      for (size_t k = 0; k < steps; k++) {
        const double c = (double)std::min((size_t)batch, m->rows - k * (size_t)batch);
        s += losses[k] * c;
        wsum += c;
      }
      epoch_loss[ep] = (float)(s / wsum);
    }
  }
  return XPBI_OK;
}

}  // extern "C"
