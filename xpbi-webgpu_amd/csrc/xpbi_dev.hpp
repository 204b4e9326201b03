// Device-side types and math of the XPBI substep (gfx950).
//
// Restates the math library of the reference, web/js/shaders/common.js:62-396,
// as HIP __device__ functions.  Layout in HBM is structure-of-arrays with
// 16-byte lanes: a 3x3 matrix per particle is three planes (float4 a, float4 b,
// float c) holding its nine column-major entries, so a wave reads each plane
// with fully coalesced dwordx4 / dword loads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xpbi {

typedef uint32_t u32;

// ---------------------------------------------------------------- records
struct DevMat {  // 64-B record, common.js:21-33
  float mu, lam, density;
  u32 kind;
  float p0, p1, p2, p3;
  float color[4];
  float boundFriction, mass, pad0, pad1;
};
static_assert(sizeof(DevMat) == 64, "material record");

enum { MAT_ELASTIC = 0, MAT_DP = 1, MAT_NACC = 2, MAT_VM = 3, MAT_HB = 4, MAT_SNOW = 5, MAT_WATER = 6 };

// Kernel parameters, passed by value (lives in SGPRs).  The first block is the
// Params uniform of common.js:35-60; the tail holds host-derived constants
// (formed in f32 exactly as the per-pair expressions of common.js would).
struct KP {
  float ox, oy, oz, cell;
  int gx, gy, gz;
  u32 C;
  float dminx, dminy, dminz, dt;
  float dmaxx, dmaxy, dmaxz, h;
  float grx, gry, grz, support;
  float cpx, cpy, cpz, cR;
  float cvx, cvy, cvz, cActive;
  u32 N;
  float V0, xsphC, omega;
  float pbfEps, sCorrK, contactOmega, kDamp;
  // derived
  float supp2;   // support * support
  float sigW;    // 21 / (16 pi h^3)            (common.js:67)
  float poly6C;  // 315 / (64 pi H^9)           (common.js:86)
  float spikyC;  // -45 / (pi H^6)              (common.js:93)
  float w0;      // kernelPoly6(0.2 support)    (simkernels.js:465)
  float poly6_0; // kernelPoly6(0)              (simkernels.js:416)
  float alphaT;  // (1/V0) / dt^2               (simkernels.js:302)
  float vmax;    // 0.45 support / dt           (simkernels.js:535)
  float dvCap;   // 0.2 support / dt            (simkernels.js:361)
  float minD;    // 0.75 h                      (simkernels.js:333)
  u32 gxy;       // gx * gy
  float invH;    // 1 / h
  float halfInvH;  // 0.5 / h
  float gradC;   // -5 sigma / h^2: grad W = dxv * gradC * (1 - q/2)^3
  float invW0;   // 1 / max(w0, 1e-12): the s_corr ratio W / W(0.2 support) as a product
  // the scene's place in a batch: its cells are cellBase .. cellBase + C - 1 of
  // the shared cell table, its particles pBase .. pEnd - 1 of the shared arrays
  u32 cellBase, pBase, pEnd;
};

// ---------------------------------------------------------------- vectors
template <class T> struct t3 { T x, y, z; };
typedef t3<float> f3;
typedef t3<double> d3;
template <class T> __device__ __forceinline__ t3<T> mk3(T a, T b, T c) { return t3<T>{a, b, c}; }
template <class T> __device__ __forceinline__ t3<T> splat3(T s) { return t3<T>{s, s, s}; }
template <class T> __device__ __forceinline__ t3<T> operator+(t3<T> a, t3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> __device__ __forceinline__ t3<T> operator-(t3<T> a, t3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> __device__ __forceinline__ t3<T> operator-(t3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> __device__ __forceinline__ t3<T> operator*(t3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> __device__ __forceinline__ t3<T> operator*(T s, t3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> __device__ __forceinline__ t3<T> operator*(t3<T> a, t3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> __device__ __forceinline__ t3<T> operator/(t3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> __device__ __forceinline__ t3<T> operator/(t3<T> a, t3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> __device__ __forceinline__ T dot(t3<T> a, t3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float xsqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ double xsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float xabs(float a) { return fabsf(a); }
__device__ __forceinline__ double xabs(double a) { return fabs(a); }
__device__ __forceinline__ float xrcp(float a) { return __builtin_amdgcn_rcpf(a); }
__device__ __forceinline__ double xrcp(double a) { return 1.0 / a; }
__device__ __forceinline__ float xrsq(float a) { return __builtin_amdgcn_rsqf(a); }
__device__ __forceinline__ double xrsq(double a) { return 1.0 / sqrt(a); }
__device__ __forceinline__ float fsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ double fsqrt(double a) { return sqrt(a); }
template <class T> __device__ __forceinline__ T len(t3<T> a) { return xsqrt(dot(a, a)); }
template <class T> __device__ __forceinline__ t3<T> normalize(t3<T> a) { return a / len(a); }
template <class T> __device__ __forceinline__ t3<T> cross(t3<T> a, t3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> __device__ __forceinline__ T wmin(T a, T b) { return b < a ? b : a; }
template <class T> __device__ __forceinline__ T wmax(T a, T b) { return a < b ? b : a; }
template <class T> __device__ __forceinline__ T wclamp(T e, T lo, T hi) { return wmin(wmax(e, lo), hi); }
__device__ __forceinline__ f3 xyz(float4 a) { return f3{a.x, a.y, a.z}; }
__device__ __forceinline__ f3 vlog(f3 e) { return {logf(e.x), logf(e.y), logf(e.z)}; }
__device__ __forceinline__ f3 vexp(f3 e) { return {expf(e.x), expf(e.y), expf(e.z)}; }
__device__ __forceinline__ f3 vsqrt(f3 e) { return {sqrtf(e.x), sqrtf(e.y), sqrtf(e.z)}; }
__device__ __forceinline__ f3 vmaxs(f3 e, float lo) { return {wmax(e.x, lo), wmax(e.y, lo), wmax(e.z, lo)}; }
__device__ __forceinline__ f3 vclamp(f3 e, float lo, float hi) {
  return {wclamp(e.x, lo, hi), wclamp(e.y, lo, hi), wclamp(e.z, lo, hi)};
}
template <class T> __device__ __forceinline__ T comp(t3<T> v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
template <class T> __device__ __forceinline__ void setc(t3<T>& v, int k, T s) {
  if (k == 0) v.x = s; else if (k == 1) v.y = s; else v.z = s;
}

// column-major 3x3 (c[col])
template <class T> struct t33 { t3<T> c[3]; };
typedef t33<float> f33;
typedef t33<double> d33;
template <class T> __device__ __forceinline__ t33<T> mzero() { return t33<T>{{splat3<T>(0), splat3<T>(0), splat3<T>(0)}}; }
template <class T> __device__ __forceinline__ t33<T> mident() {
  return t33<T>{{mk3<T>(1, 0, 0), mk3<T>(0, 1, 0), mk3<T>(0, 0, 1)}};
}
template <class T> __device__ __forceinline__ t3<T> mul(const t33<T>& A, t3<T> v) {
  return A.c[0] * v.x + A.c[1] * v.y + A.c[2] * v.z;
}
template <class T> __device__ __forceinline__ t33<T> mul(const t33<T>& A, const t33<T>& B) {
  return t33<T>{{mul(A, B.c[0]), mul(A, B.c[1]), mul(A, B.c[2])}};
}
template <class T> __device__ __forceinline__ t33<T> transpose(const t33<T>& A) {
  return t33<T>{{mk3(A.c[0].x, A.c[1].x, A.c[2].x), mk3(A.c[0].y, A.c[1].y, A.c[2].y),
                 mk3(A.c[0].z, A.c[1].z, A.c[2].z)}};
}
template <class T> __device__ __forceinline__ t33<T> madd(const t33<T>& A, const t33<T>& B) {
  return t33<T>{{A.c[0] + B.c[0], A.c[1] + B.c[1], A.c[2] + B.c[2]}};
}
template <class T> __device__ __forceinline__ t33<T> mscale(const t33<T>& A, T s) {
  return t33<T>{{A.c[0] * s, A.c[1] * s, A.c[2] * s}};
}
template <class T> __device__ __forceinline__ t33<T> outer(t3<T> a, t3<T> b) {
  return t33<T>{{a * b.x, a * b.y, a * b.z}};
}
template <class T> __device__ __forceinline__ T det3(const t33<T>& m) {
  return m.c[0].x * (m.c[1].y * m.c[2].z - m.c[2].y * m.c[1].z) -
         m.c[1].x * (m.c[0].y * m.c[2].z - m.c[2].y * m.c[0].z) +
         m.c[2].x * (m.c[0].y * m.c[1].z - m.c[1].y * m.c[0].z);
}
template <class T, class S> __device__ __forceinline__ t33<T> mcast(const t33<S>& a) {
  t33<T> r;
  for (int j = 0; j < 3; j++) r.c[j] = mk3<T>(T(a.c[j].x), T(a.c[j].y), T(a.c[j].z));
  return r;
}
template <class T, class S> __device__ __forceinline__ t3<T> vcast(t3<S> a) { return mk3<T>(T(a.x), T(a.y), T(a.z)); }

// ---------------------------------------------------------------- planes
struct M3P {  // nine floats per particle: a = (c0.xyz, c1.x), b = (c1.yz, c2.xy), c = c2.z
  float4* a;
  float4* b;
  float* c;
};
struct M3C {
  const float4* a;
  const float4* b;
  const float* c;
};
__device__ __forceinline__ f33 ldm(const float4* __restrict__ A, const float4* __restrict__ B,
                                   const float* __restrict__ Cc, u32 i) {
  float4 a = A[i], b = B[i];
  float c = Cc[i];
  return f33{{f3{a.x, a.y, a.z}, f3{a.w, b.x, b.y}, f3{b.z, b.w, c}}};
}
__device__ __forceinline__ f33 ldm(M3C m, u32 i) { return ldm(m.a, m.b, m.c, i); }
__device__ __forceinline__ f33 ldm(M3P m, u32 i) { return ldm(m.a, m.b, m.c, i); }
__device__ __forceinline__ void stm(M3P m, u32 i, const f33& v) {
  m.a[i] = make_float4(v.c[0].x, v.c[0].y, v.c[0].z, v.c[1].x);
  m.b[i] = make_float4(v.c[1].y, v.c[1].z, v.c[2].x, v.c[2].y);
  m.c[i] = v.c[2].z;
}
// Sp = dlam GL of solveA (and the unscaled GL of the warm-start features): a and
// b as in M3P, cx = (x_i, the ninth float), so that the all-solid gather of
// solveB reads a neighbour's position and its Sp as three 16-byte records
struct SpP {
  float4* a;
  float4* b;
  float4* cx;
};
struct SpC {
  const float4* a;
  const float4* b;
  const float4* cx;
};
__device__ __forceinline__ void stsp(SpP m, u32 i, const f33& v, float4 pi) {
  m.a[i] = make_float4(v.c[0].x, v.c[0].y, v.c[0].z, v.c[1].x);
  m.b[i] = make_float4(v.c[1].y, v.c[1].z, v.c[2].x, v.c[2].y);
  m.cx[i] = make_float4(pi.x, pi.y, pi.z, v.c[2].z);
}

// ---------------------------------------------------------------- SPH kernels
// common.js:63-69
__device__ __forceinline__ float kernelW(const KP& P, float r) {
  float q = r / P.h;
  if (q >= 2.0f) return 0.0f;
  float t = 1.0f - 0.5f * q;
  return P.sigW * t * t * t * t * (2.0f * q + 1.0f);
}
// common.js:71-79, dxv = xp - xb
__device__ __forceinline__ f3 kernelGradW(const KP& P, f3 dxv) {
  float r = len(dxv);
  float q = r / P.h;
  if (q >= 2.0f || r < 1e-9f) return splat3(0.0f);
  float t = 1.0f - 0.5f * q;
  float dwdq = P.sigW * (-5.0f * q) * t * t * t;
  return dxv * (dwdq / (P.h * r));
}
// Same kernels in closed form for the hot sweeps.  With q = r/h and t = 1 - q/2:
//   dwdq / (h r) = sigma (-5 q) t^3 / (h r) = (-5 sigma / h^2) t^3,
// so grad W = dxv * gradC * t^3 needs one sqrt and no division.  r2 = |dxv|^2 is
// already known (< support^2) from the neighbour test; the r < 1e-9 case of the
// reference returns zero.
// These kernels take the scene record KP or its register copy SphC (the same
// field names): read through the record inside a neighbour loop, the constants
// are reloaded from memory at every neighbour where scalar registers run short.
template <class K> __device__ __forceinline__ f3 gradWfast(const K& P, f3 dxv, float r2) {
  float r = fsqrt(r2);
  float t = 1.0f - r * P.halfInvH;
  float c = (r2 < 1e-18f) ? 0.0f : P.gradC * (t * t * t);
  return dxv * c;
}
template <class K> __device__ __forceinline__ float kernelWfast(const K& P, float r2) {
  float q = fsqrt(r2) * P.invH;
  float t = 1.0f - 0.5f * q;
  float t2 = t * t;
  return P.sigW * (t2 * t2) * (2.0f * q + 1.0f);
}
// the scene's SPH constants of the neighbour sweeps, copied into registers
struct SphC {
  float halfInvH, gradC, invH, sigW, support, poly6C, spikyC, invW0, sCorrK, dt, minD;
};
__device__ __forceinline__ SphC sphConsts(const KP& P) {
  return SphC{P.halfInvH, P.gradC, P.invH, P.sigW, P.support, P.poly6C, P.spikyC, P.invW0, P.sCorrK, P.dt, P.minD};
}

// common.js:82-87
template <class K> __device__ __forceinline__ float kernelPoly6(const K& P, float r) {
  float H = P.support;
  if (r >= H) return 0.0f;
  float x = H * H - r * r;
  return P.poly6C * x * x * x;
}
// common.js:88-94 with the pair distance r = |dxv| already formed and the
// division by r as a product with the hardware reciprocal (1 ulp)
template <class K> __device__ __forceinline__ f3 kernelSpikyGradR(const K& P, f3 dxv, float r) {
  float H = P.support;
  if (r >= H || r < 1e-9f) return splat3(0.0f);
  float t = H - r;
  return dxv * (P.spikyC * t * t * __builtin_amdgcn_rcpf(r));
}

// ---------------------------------------------------------------- grid hash
// common.js:97-103.  IEEE f32 division (hipcc's correctly rounded default) and
// floor; v_cvt_i32_f32 saturates and maps NaN to 0 like the WGSL conversion.
struct i3 { int x, y, z; };

// element idx of an array read by a gather, addressed by a 32-bit byte offset
// from the array's base (idx * sizeof(T) < 2^32: at most 2^28 16-B records), so
// that the load takes the base in scalar registers and one 32-bit vector offset
template <class T>
__device__ __forceinline__ T gat(const T* __restrict__ base, u32 idx) {
  return *(const T*)((const char*)base + idx * (u32)sizeof(T));
}
// the w lane of float4 record idx
__device__ __forceinline__ float gatW(const float4* __restrict__ base, u32 idx) {
  return *(const float*)((const char*)base + idx * 16u + 12u);
}
__device__ __forceinline__ int f2i_sat(float f) { return __float2int_rz(f); }
__device__ __forceinline__ i3 cellCoord(const KP& P, f3 p) {
  int cx = f2i_sat(floorf((p.x - P.ox) / P.cell));
  int cy = f2i_sat(floorf((p.y - P.oy) / P.cell));
  int cz = f2i_sat(floorf((p.z - P.oz) / P.cell));
  cx = min(max(cx, 0), P.gx - 1);
  cy = min(max(cy, 0), P.gy - 1);
  cz = min(max(cz, 0), P.gz - 1);
  return i3{cx, cy, cz};
}
__device__ __forceinline__ u32 cellIndex(const KP& P, i3 c) {
  return P.cellBase + u32(c.x) + u32(c.y) * u32(P.gx) + u32(c.z) * P.gxy;
}
// pos.w holds the material index as a float value (simkernels.js:41-44); the
// mask keeps a corrupt tag inside the 8-entry material table.
__device__ __forceinline__ u32 matOfTag(float w) { return __float2uint_rz(w) & 7u; }

// ---------------------------------------------------------------- svd3
template <class T> struct SVD3 {
  t33<T> U;
  t3<T> sig;
  t33<T> V;
  bool degenerate;  // a fallback branch fired or sigma_min << sigma_max
};
template <class T> __device__ __forceinline__ T& el(t33<T>& m, int col, int row) {
  return row == 0 ? m.c[col].x : (row == 1 ? m.c[col].y : m.c[col].z);
}
template <class T> __device__ __forceinline__ T elc(const t33<T>& m, int col, int row) {
  return row == 0 ? m.c[col].x : (row == 1 ? m.c[col].y : m.c[col].z);
}
// Symmetric 3x3 held as its 6 distinct entries.
template <class T> struct Sym3 { T a00, a11, a22, a01, a02, a12; };

// common.js:108-143: one cyclic-Jacobi rotation in the (p,q) plane, S <- J^T S J,
// V <- V J, with the same tau/t/c/s as the reference.  S is symmetric, so only
// its 6 distinct entries are rotated and the annihilated S_pq is set to 0
// (the reference recomputes it as rounding noise).
template <class T>
__device__ __forceinline__ void jrot(T& Spp, T& Sqq, T& Spq, T& Srp, T& Srq, t3<T>& vp, t3<T>& vq) {
  if (xabs(Spq) < T(1e-12)) return;
  T tau = (Sqq - Spp) * xrcp(T(2) * Spq);
  T rt = xsqrt(T(1) + tau * tau);
  T t = xrcp(tau >= T(0) ? tau + rt : tau - rt);
  T c = xrsq(T(1) + t * t);
  T s = t * c;
  T tq = t * Spq;
  Spp = Spp - tq;
  Sqq = Sqq + tq;
  Spq = T(0);
  T rp = Srp, rq = Srq;
  Srp = c * rp - s * rq;
  Srq = s * rp + c * rq;
  t3<T> a = vp, b = vq;
  vp = c * a - s * b;
  vq = s * a + c * b;
}
template <class T> __device__ __forceinline__ void swp(T& a, T& b) { T t = a; a = b; b = t; }
template <class T> struct JTol;
template <> struct JTol<float> { static constexpr float v = 1e-7f; };
template <> struct JTol<double> { static constexpr double v = 1e-15; };
// common.js:145-180.  The reference runs 6 fixed sweeps; sweeps stop early here
// once the off-diagonal mass is below the type's rounding level (a further
// sweep would only move rounding noise).
template <class T> __device__ __forceinline__ SVD3<T> svd3(const t33<T>& F) {
  const t3<T> f0 = F.c[0], f1 = F.c[1], f2 = F.c[2];
  T a00 = dot(f0, f0), a11 = dot(f1, f1), a22 = dot(f2, f2);
  T a01 = dot(f0, f1), a02 = dot(f0, f2), a12 = dot(f1, f2);
  t3<T> w0 = mk3<T>(1, 0, 0), w1 = mk3<T>(0, 1, 0), w2 = mk3<T>(0, 0, 1);
#pragma unroll 1
  for (int sweep = 0; sweep < 6; sweep++) {
    jrot(a00, a11, a01, a02, a12, w0, w1);  // (0,1): third index 2
    jrot(a00, a22, a02, a01, a12, w0, w2);  // (0,2): third index 1
    jrot(a11, a22, a12, a01, a02, w1, w2);  // (1,2): third index 0
    T off = xabs(a01) + xabs(a02) + xabs(a12);
    if (off <= JTol<T>::v * (a00 + a11 + a22)) break;
  }
  t3<T> sig = mk3(xsqrt(wmax(a00, T(0))), xsqrt(wmax(a11, T(0))), xsqrt(wmax(a22, T(0))));
  t3<T> v0 = w0, v1 = w1, v2 = w2;
  if (sig.x < sig.y) { swp(sig.x, sig.y); swp(v0, v1); }
  if (sig.x < sig.z) { swp(sig.x, sig.z); swp(v0, v2); }
  if (sig.y < sig.z) { swp(sig.y, sig.z); swp(v1, v2); }
  if (dot(cross(v0, v1), v2) < T(0)) v2 = -v2;
  bool degen = false;
  t3<T> u0, u1;
  const T eps = T(1e-8);
  if (sig.x > eps) {
    t3<T> Fv0 = mul(F, v0);
    u0 = Fv0 * xrsq(dot(Fv0, Fv0));
  } else { u0 = mk3<T>(1, 0, 0); degen = true; }
  if (sig.y > eps) {
    t3<T> Fv1 = mul(F, v1);
    u1 = Fv1 - dot(Fv1, u0) * u0;
    T l2 = dot(u1, u1);
    if (l2 > eps * eps) u1 = u1 * xrsq(l2);
    else { u1 = normalize(cross(mk3<T>(0, 1, 0), u0) + splat3(T(1e-4))); degen = true; }
  } else {
    t3<T> a = cross(u0, mk3<T>(0, 1, 0));
    if (len(a) < T(1e-4)) a = cross(u0, mk3<T>(1, 0, 0));
    u1 = normalize(a);
    degen = true;
  }
  t3<T> u2 = cross(u0, u1);
  if (det3(F) < T(0)) sig.z = -sig.z;
  SVD3<T> r;
  r.U = t33<T>{{u0, u1, u2}};
  r.sig = sig;
  r.V = t33<T>{{v0, v1, v2}};
  r.degenerate = degen;
  return r;
}
// f32 SVD of F^T F loses sigma_min when sigma_min/sigma_max is small (the
// condition number is squared).  That near-degenerate branch is recomputed in
// FP64; everything else stays FP32.
__device__ __forceinline__ SVD3<float> svd3_guarded(const f33& F) {
  SVD3<float> s = svd3<float>(F);
  float smin = fabsf(s.sig.z);
  if (s.degenerate || smin < 0.02f * s.sig.x) {
    SVD3<double> d = svd3<double>(mcast<double>(F));
    s.U = mcast<float>(d.U);
    s.V = mcast<float>(d.V);
    s.sig = vcast<float>(d.sig);
  }
  return s;
}
// common.js:182-184  U diag(d) V^T
template <class T> __device__ __forceinline__ t33<T> diagMul(const t33<T>& U, t3<T> d, const t33<T>& Vm) {
  t33<T> A{{U.c[0] * d.x, U.c[1] * d.y, U.c[2] * d.z}};
  return mul(A, transpose(Vm));
}

// ---------------------------------------------------------------- StVK-Hencky
// common.js:187-201
__device__ __forceinline__ f3 safeSig(f3 s) { return vclamp(s, 1e-3f, 1e3f); }
__device__ __forceinline__ float stvkPsi(f3 sig, float mu, float lam) {
  f3 e = vlog(safeSig(sig));
  float tr = e.x + e.y + e.z;
  return mu * dot(e, e) + 0.5f * lam * tr * tr;
}
__device__ __forceinline__ f3 stvkDPsi(f3 sig, float mu, float lam) {
  f3 s = safeSig(sig);
  f3 e = vlog(s);
  float tr = e.x + e.y + e.z;
  return (2.0f * mu * e + splat3(lam * tr)) / s;
}

// ---------------------------------------------------------------- return maps
struct MatV {  // material in registers
  float mu, lam, density;
  u32 kind;
  float p0, p1, p2, p3;
  float fric, mass;
};
__device__ __forceinline__ MatV loadMat(const DevMat* m, u32 k) {
  const DevMat& r = m[k];
  return MatV{r.mu, r.lam, r.density, r.kind, r.p0, r.p1, r.p2, r.p3, r.boundFriction, r.mass};
}
// common.js:206-221
__device__ __forceinline__ f3 rmDruckerPrager(f3 sig, const MatV& m) {
  float alpha = m.p0, coh = m.p1;
  f3 e = vlog(safeSig(sig)) - splat3(coh);
  float tr = e.x + e.y + e.z;
  f3 dev = e - splat3(tr / 3.0f);
  float dn = len(dev);
  if (dn < 1e-9f || tr > 0.0f) return vexp(splat3(coh));
  float dg = dn + alpha * (3.0f * m.lam + 2.0f * m.mu) / (2.0f * m.mu) * tr;
  if (dg <= 0.0f) return sig;
  f3 en = e - (dg / dn) * dev;
  return vexp(en + splat3(coh));
}
// common.js:224-234
__device__ __forceinline__ f3 rmVonMises(f3 sig, const MatV& m) {
  float sigY = m.p0;
  f3 e = vlog(safeSig(sig));
  float tr = e.x + e.y + e.z;
  f3 dev = e - splat3(tr / 3.0f);
  float dn = len(dev);
  float dg = dn - sigY / (2.0f * m.mu);
  if (dg <= 0.0f || dn < 1e-9f) return sig;
  f3 en = e - (dg / dn) * dev;
  return vexp(en);
}
// common.js:244-280
//
// hint in [0, 1] (learned warm start) is a predicted root position t in
// s = sY + t (sPre - sY): the bisection then starts from [t - 1/64, t + 1/64]
// when g changes sign across it (13 halvings reach the 2^-18 width of the
// reference's 18 over the full bracket), else from the full bracket.  *tOut
// receives the root position of the bisection (NaN when none ran).
__device__ __forceinline__ f3 rmHerschelBulkley(f3 sig, const MatV& m, float dt, float hint = -1.0f,
                                                float* tOut = nullptr) {
  float sigY = m.p0, hexp = m.p1, eta = m.p2;
  if (tOut) *tOut = __builtin_nanf("");
  f3 s3 = safeSig(sig);
  float J = s3.x * s3.y * s3.z;
  f3 bbar = s3 * s3 * powf(J, -2.0f / 3.0f);
  float trb = bbar.x + bbar.y + bbar.z;
  f3 devb = bbar - splat3(trb / 3.0f);
  float sPre = m.mu * len(devb);
  float sY = 0.816496580927726f * sigY;  // sqrt(2/3)
  if (sPre <= sY || sPre < 1e-12f) return sig;
  float muT = (trb / 3.0f) * m.mu;
  float sNew;
  if (fabsf(hexp - 1.0f) < 1e-3f || eta < 1e-9f) {
    sNew = sPre - (sPre - sY) / (1.0f + eta / (2.0f * muT * dt));
  } else {
    float ih = 1.0f / hexp;
    float ei = powf(eta, ih);
    auto g = [&](float x) { return ei * (x - sPre) + 2.0f * muT * dt * powf(wmax(x - sY, 0.0f), ih); };
    float lo = sY, hi = sPre;
    int steps = 18;
    if (hint >= 0.0f && hint <= 1.0f) {
      const float a = wmax(hint - 1.0f / 64.0f, 0.0f), b = wmin(hint + 1.0f / 64.0f, 1.0f);
      const float lo2 = sY + a * (sPre - sY), hi2 = sY + b * (sPre - sY);
      if ((a == 0.0f || !(g(lo2) > 0.0f)) && (b == 1.0f || g(hi2) > 0.0f)) {
        lo = lo2;
        hi = hi2;
        steps = 13;
      }
    }
#pragma unroll 1
    for (int it = 0; it < steps; it++) {
      float mid = 0.5f * (lo + hi);
      if (g(mid) > 0.0f) hi = mid; else lo = mid;
    }
    sNew = 0.5f * (lo + hi);
    if (tOut) *tOut = (sNew - sY) / (sPre - sY);
  }
  f3 bNew = devb * (sNew / sPre) + splat3(trb / 3.0f);
  bNew = vmaxs(bNew, 1e-6f);
  float detB = bNew.x * bNew.y * bNew.z;
  bNew = bNew * powf(detB, -1.0f / 3.0f);
  return vsqrt(bNew) * powf(J, 1.0f / 3.0f);
}
// common.js:283-285
__device__ __forceinline__ f3 rmSnowClamp(f3 sig, const MatV& m) {
  float lo = 1.0f - m.p0, hi = 1.0f + m.p1;
  return f3{wclamp(sig.x, lo, hi), wclamp(sig.y, lo, hi), wclamp(sig.z, lo, hi)};
}
// common.js:293-364
__device__ __noinline__ f3 rmNACC(f3 sig, const MatV m, float alphaIn, float* alphaOut) {
  float beta = m.p1, xi = m.p2, Mf = m.p3;
  float mu = m.mu;
  float kappa = (2.0f / 3.0f) * mu + m.lam;
  float alpha = alphaIn;
  f3 s3 = safeSig(sig);
  float Jtr = s3.x * s3.y * s3.z;
  f3 b2 = s3 * s3;
  float trb = b2.x + b2.y + b2.z;
  f3 devb = b2 - splat3(trb / 3.0f);
  f3 sHat = mu * powf(Jtr, -2.0f / 3.0f) * devb;
  float sNorm = len(sHat);
  const float sq15 = 1.2247448713915890f;  // sqrt(1.5)
  float q = sq15 * sNorm;
  float p = -0.5f * kappa * (Jtr * Jtr - 1.0f);
  alpha = wclamp(alpha, -0.25f, 0.4f);
  float p0 = kappa * (0.00001f + sinhf(wmin(xi * wmax(-alpha, 0.0f), 4.0f)));
  float b = 1.0f + 2.0f * beta;
  f3 outSig = sig;
  if (p > p0) {
    float Jn = sqrtf(wmax(-2.0f * p0 / kappa + 1.0f, 1e-6f));
    outSig = splat3(powf(Jn, 1.0f / 3.0f));
    alpha = alpha + logf(Jtr / Jn);
  } else if (p < -beta * p0) {
    float Jn = sqrtf(wmax(2.0f * beta * p0 / kappa + 1.0f, 1e-6f));
    outSig = splat3(powf(Jn, 1.0f / 3.0f));
    alpha = alpha + logf(Jtr / Jn);
  } else {
    float y = b * q * q + Mf * Mf * (p + beta * p0) * (p - p0);
    if (y > 1e-4f && sNorm > 1e-9f) {
      if (p0 > 1e-4f && p < p0 - 1e-4f && p > -beta * p0 + 1e-4f) {
        float pc = (1.0f - beta) * p0 * 0.5f;
        float dpc = pc - p, dq = -q;
        float lenDir = sqrtf(dpc * dpc + dq * dq);
        if (lenDir > 1e-9f) {
          float Dp = (pc - p) / lenDir;
          float Dq = -q / lenDir;
          float Aq = Mf * Mf * Dp * Dp + b * Dq * Dq;
          float Bq = Mf * Mf * Dp * (2.0f * pc - p0 + beta * p0);
          float Cq = Mf * Mf * (pc + beta * p0) * (pc - p0);
          float disc = Bq * Bq - 4.0f * Aq * Cq;
          if (disc > 0.0f && fabsf(Aq) > 1e-12f) {
            float l1 = (-Bq + sqrtf(disc)) / (2.0f * Aq);
            float l2 = (-Bq - sqrtf(disc)) / (2.0f * Aq);
            float px1 = pc + l1 * Dp;
            float px2 = pc + l2 * Dp;
            float px = px1;
            if ((p - pc) * (px2 - pc) > 0.0f) px = px2;
            float Jx2 = -2.0f * px / kappa + 1.0f;
            if (Jx2 > 1e-8f) {
              float Jx = sqrtf(Jx2);
              if (Jx > 1e-4f) alpha = alpha + logf(Jtr / Jx);
            }
          }
        }
      }
      float qN = Mf * sqrtf(wmax(-(p + beta * p0) * (p - p0), 0.0f) / b);
      float sNewNorm = qN / sq15;
      f3 bNew = (powf(Jtr, 2.0f / 3.0f) / mu) * (sNewNorm / sNorm) * sHat + splat3(trb / 3.0f);
      bNew = vmaxs(bNew, 1e-6f);
      outSig = vsqrt(bNew);
    }
  }
  *alphaOut = alpha;
  return outSig;
}
// common.js:367-375
__device__ __forceinline__ MatV hardenedMat(MatV m, float state) {
  if (m.kind == MAT_SNOW) {
    float s = expf(wmin(m.p2 * (1.0f - state), 5.0f));
    m.mu = m.mu * s;
    m.lam = m.lam * s;
  }
  return m;
}
// common.js:379-396
// hbHint / hbT: the HB bracket hint and root position (rmHerschelBulkley)
__device__ __forceinline__ f3 returnMap(f3 sig, const MatV& m, float state, float dt, float* stateOut,
                                        float hbHint = -1.0f, float* hbT = nullptr) {
  *stateOut = state;
  switch (m.kind) {
    case MAT_DP: return rmDruckerPrager(sig, m);
    case MAT_NACC: return rmNACC(sig, m, state, stateOut);
    case MAT_VM: return rmVonMises(sig, m);
    case MAT_HB: return rmHerschelBulkley(sig, m, dt, hbHint, hbT);
    case MAT_SNOW: {
      MatV mm = hardenedMat(m, state);
      f3 sNew = rmSnowClamp(sig, mm);
      float jOld = sig.x * sig.y * sig.z;
      float jNew = sNew.x * sNew.y * sNew.z;
      *stateOut = wclamp(state * jOld / wmax(jNew, 1e-6f), 0.2f, 2.0f);
      return sNew;
    }
    default: return sig;
  }
}

// ---------------------------------------------------------------- launch helpers
// Blocks are dispatched round-robin over the 8 XCDs; remap so that each XCD
// walks one contiguous run of particles (cell order), keeping the neighbour
// rows of a block in that XCD's own L2.
__device__ __forceinline__ u32 xcdBlock() {
  u32 nb = gridDim.x, b = blockIdx.x;
  u32 q = nb >> 3, r = nb & 7u, x = b & 7u, k = b >> 3;
  return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// ---------------------------------------------------------------- work table
// A launch walks a table of block-sized particle runs that never cross a scene
// boundary: entry t = (scene, first particle), the run ending at
// min(first + 256, scene end).  A batch of scenes is one launch per kernel.
struct WorkTab {
  const uint2* __restrict__ tab;
  u32 nb;
};
// fn(scene, first) for this block's entries.  kPersist = false: one entry per
// block (the grid covers the table), XCD-contiguous.  kPersist = true: a
// persistent grid (a multiple of 8 blocks) where XCD x walks the x-th
// contiguous eighth of the table; a separate instantiation, because the loop
// costs registers the one-entry form does not need.
template <bool kPersist, class Fn>
__device__ __forceinline__ void forWork(const WorkTab& W, Fn&& fn) {
  if (!kPersist) {
    const u32 t = xcdBlock();
    if (t < W.nb) fn(W.tab[t].x, W.tab[t].y);
    return;
  }
  const u32 x = blockIdx.x & 7u, per = (W.nb + 7u) >> 3, step = gridDim.x >> 3;
  const u32 t1 = min(x * per + per, W.nb);
  for (u32 t = x * per + (blockIdx.x >> 3); t < t1; t += step) fn(W.tab[t].x, W.tab[t].y);
}

}  // namespace xpbi
