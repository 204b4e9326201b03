// CPU oracle for the XPBI substep -- TEST INFRASTRUCTURE ONLY.
//
// A line-by-line C++ restatement of the reference WGSL substep, templated on
// the arithmetic type T (float reproduces the reference's f32 arithmetic, double
// gives the FP64 error yardstick).  Only the tests, the smoke check and the CPU
// baseline leg of the benchmark may load it; the product library never does.
//
// Every function cites the reference file:line it restates
// (paths relative to the reference root):
//   web/js/shaders/common.js     math library (kernels, hash, svd3, StVK, return maps)
//   web/js/shaders/simkernels.js compute entry points
//   web/js/sim.js                dispatch schedule and buffer layout
//
// Semantics kept from WGSL: column-major mat3x3 (c[col].row), f32 arithmetic in
// source order, select()/clamp() definitions, u32(f32)/i32(f32) truncation with
// clamping.  Kept reference quirks: the 384-candidate row cap, DP's tip on zero
// deviator, double snow hardening inside returnMap, stale contacts in solveB,
// velIn/velOut ping-pong and its parity copy.  The one deliberate difference:
// the within-cell order of the counting sort is STABLE (ascending original
// index), one of the orders the reference's atomicSub scatter can produce.
//
// Pinning status (what checks this file independently of itself) is listed in
// the header comment of oracle/xpbi_oracle.cpp.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>

namespace xo {

using u32 = uint32_t;

template <class T> struct v3 {
  T x, y, z;
  T& operator[](int k) { return k == 0 ? x : (k == 1 ? y : z); }
  T operator[](int k) const { return k == 0 ? x : (k == 1 ? y : z); }
};
template <class T> inline v3<T> V(T a, T b, T c) { return v3<T>{a, b, c}; }
template <class T> inline v3<T> operator+(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline v3<T> operator-(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline v3<T> operator-(v3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline v3<T> operator*(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline v3<T> operator*(T s, v3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline v3<T> operator*(v3<T> a, v3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline v3<T> operator/(v3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline v3<T> operator/(v3<T> a, v3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> inline T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T length(v3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline v3<T> normalize(v3<T> a) { return a / length(a); }
template <class T> inline v3<T> cross(v3<T> a, v3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T wmin(T a, T b) { return b < a ? b : a; }  // WGSL min (NaN-agnostic)
template <class T> inline T wmax(T a, T b) { return a < b ? b : a; }
template <class T> inline T wclamp(T e, T lo, T hi) { return wmin(wmax(e, lo), hi); }
template <class T> inline v3<T> vclamp(v3<T> e, T lo, T hi) {
  return {wclamp(e.x, lo, hi), wclamp(e.y, lo, hi), wclamp(e.z, lo, hi)};
}
template <class T> inline v3<T> vmax(v3<T> e, T lo) { return {wmax(e.x, lo), wmax(e.y, lo), wmax(e.z, lo)}; }
template <class T> inline v3<T> vlog(v3<T> e) { return {std::log(e.x), std::log(e.y), std::log(e.z)}; }
template <class T> inline v3<T> vexp(v3<T> e) { return {std::exp(e.x), std::exp(e.y), std::exp(e.z)}; }
template <class T> inline v3<T> vsqrt(v3<T> e) { return {std::sqrt(e.x), std::sqrt(e.y), std::sqrt(e.z)}; }
template <class T> inline v3<T> splat(T s) { return {s, s, s}; }

// column-major 3x3: c[col] is a column vector, element (row r, col j) = c[j][r]
template <class T> struct m3 { v3<T> c[3]; };
template <class T> inline m3<T> mzero() { return m3<T>{{splat<T>(0), splat<T>(0), splat<T>(0)}}; }
template <class T> inline m3<T> mident() {
  return m3<T>{{V<T>(1, 0, 0), V<T>(0, 1, 0), V<T>(0, 0, 1)}};
}
template <class T> inline v3<T> mul(const m3<T>& A, v3<T> v) {
  return A.c[0] * v.x + A.c[1] * v.y + A.c[2] * v.z;
}
template <class T> inline m3<T> mul(const m3<T>& A, const m3<T>& B) {
  return m3<T>{{mul(A, B.c[0]), mul(A, B.c[1]), mul(A, B.c[2])}};
}
template <class T> inline m3<T> transpose(const m3<T>& A) {
  return m3<T>{{V(A.c[0].x, A.c[1].x, A.c[2].x), V(A.c[0].y, A.c[1].y, A.c[2].y),
                V(A.c[0].z, A.c[1].z, A.c[2].z)}};
}
template <class T> inline m3<T> madd(const m3<T>& A, const m3<T>& B) {
  return m3<T>{{A.c[0] + B.c[0], A.c[1] + B.c[1], A.c[2] + B.c[2]}};
}
template <class T> inline m3<T> mscale(const m3<T>& A, T s) {
  return m3<T>{{A.c[0] * s, A.c[1] * s, A.c[2] * s}};
}
// cofactor expansion down the first column
template <class T> inline T det(const m3<T>& m) {
  return m.c[0].x * (m.c[1].y * m.c[2].z - m.c[2].y * m.c[1].z) -
         m.c[1].x * (m.c[0].y * m.c[2].z - m.c[2].y * m.c[0].z) +
         m.c[2].x * (m.c[0].y * m.c[1].z - m.c[1].y * m.c[0].z);
}
// mat3x3(a*b.x, a*b.y, a*b.z): outer product a b^T
template <class T> inline m3<T> outer(v3<T> a, v3<T> b) { return m3<T>{{a * b.x, a * b.y, a * b.z}}; }

// ---------------------------------------------------------------- records
// 64-B material record (common.js:21-33, packed by sim.js:84-99)
struct MatRec {
  float mu, lam, density;
  u32 kind;
  float p0, p1, p2, p3;
  float color[4];
  float boundFriction, mass, pad0, pad1;
};
static_assert(sizeof(MatRec) == 64, "material record is 64 bytes");
// 256-B parameter record (common.js:35-60, packed by sim.js:159-186)
struct ParamRec {
  float gridOrigin[3];
  float cellSize;
  int32_t gridDims[3];
  u32 numCells;
  float domainMin[3];
  float dt;
  float domainMax[3];
  float h;
  float gravity[3];
  float support;
  float colliderPos[3];
  float colliderR;
  float colliderVel[3];
  float colliderActive;
  u32 N;
  float V0, xsphC, omega, pbfEps, sCorrK, time, contactOmega, kDamp;
  float pad[27];
};
static_assert(sizeof(ParamRec) == 256, "param record is 256 bytes");

enum { MAT_ELASTIC = 0, MAT_DP = 1, MAT_NACC = 2, MAT_VM = 3, MAT_HB = 4, MAT_SNOW = 5, MAT_WATER = 6 };

template <class T> struct Mat {
  T mu, lam, density;
  u32 kind;
  T p0, p1, p2, p3;
  T boundFriction, mass;
};
template <class T> inline Mat<T> toMat(const MatRec& r) {
  return Mat<T>{T(r.mu), T(r.lam), T(r.density), r.kind, T(r.p0), T(r.p1), T(r.p2), T(r.p3),
                T(r.boundFriction), T(r.mass)};
}

template <class T> struct Prm {
  v3<T> gridOrigin;
  T cellSize;
  int32_t dims[3];
  u32 numCells;
  v3<T> domainMin, domainMax, gravity, colliderPos, colliderVel;
  T dt, h, support, colliderR, colliderActive;
  u32 N;
  T V0, xsphC, omega, pbfEps, sCorrK, contactOmega, kDamp;
  // grid hash is always evaluated in f32 so that the f64 oracle shares the
  // permutation and candidate sets of the f32 one
  float fOrigin[3], fCell;
};
template <class T> inline Prm<T> toPrm(const ParamRec& r) {
  Prm<T> p;
  p.gridOrigin = V<T>(r.gridOrigin[0], r.gridOrigin[1], r.gridOrigin[2]);
  p.cellSize = T(r.cellSize);
  for (int k = 0; k < 3; k++) p.dims[k] = r.gridDims[k];
  p.numCells = r.numCells;
  p.domainMin = V<T>(r.domainMin[0], r.domainMin[1], r.domainMin[2]);
  p.domainMax = V<T>(r.domainMax[0], r.domainMax[1], r.domainMax[2]);
  p.gravity = V<T>(r.gravity[0], r.gravity[1], r.gravity[2]);
  p.colliderPos = V<T>(r.colliderPos[0], r.colliderPos[1], r.colliderPos[2]);
  p.colliderVel = V<T>(r.colliderVel[0], r.colliderVel[1], r.colliderVel[2]);
  p.dt = T(r.dt); p.h = T(r.h); p.support = T(r.support); p.colliderR = T(r.colliderR);
  p.colliderActive = T(r.colliderActive);
  p.N = r.N;
  p.V0 = T(r.V0); p.xsphC = T(r.xsphC); p.omega = T(r.omega); p.pbfEps = T(r.pbfEps);
  p.sCorrK = T(r.sCorrK); p.contactOmega = T(r.contactOmega); p.kDamp = T(r.kDamp);
  for (int k = 0; k < 3; k++) p.fOrigin[k] = r.gridOrigin[k];
  p.fCell = r.cellSize;
  return p;
}

// ---------------------------------------------------------------- SPH kernels
// common.js:62-69  Wendland C2, support 2h.  (16*pi) is a literal-only product,
// folded at compile time in WGSL, so it is formed in double and rounded once.
template <class T> inline T sigmaW(T h) { return T(21.0) / (T(16.0 * 3.14159265) * h * h * h); }
template <class T> inline T kernelW(const Prm<T>& P, T r) {
  T q = r / P.h;
  if (q >= T(2)) return T(0);
  T t = T(1) - T(0.5) * q;
  T sig = sigmaW(P.h);
  return sig * t * t * t * t * (T(2) * q + T(1));
}
// common.js:70-79  grad_{xp} W(|xp-xb|), dxv = xp - xb
template <class T> inline v3<T> kernelGradW(const Prm<T>& P, v3<T> dxv) {
  T r = length(dxv);
  T q = r / P.h;
  if (q >= T(2) || r < T(1e-9)) return splat<T>(0);
  T t = T(1) - T(0.5) * q;
  T sig = sigmaW(P.h);
  T dwdq = sig * (T(-5) * q) * t * t * t;
  return dxv * (dwdq / (P.h * r));
}
// common.js:81-87
template <class T> inline T kernelPoly6(const Prm<T>& P, T r) {
  T H = P.support;
  if (r >= H) return T(0);
  T x = H * H - r * r;
  return T(315.0) / (T(64.0 * 3.14159265) * std::pow(H, T(9))) * x * x * x;
}
// common.js:88-94
template <class T> inline v3<T> kernelSpikyGrad(const Prm<T>& P, v3<T> dxv) {
  T H = P.support;
  T r = length(dxv);
  if (r >= H || r < T(1e-9)) return splat<T>(0);
  T t = H - r;
  return dxv * (T(-45.0) / (T(3.14159265) * std::pow(H, T(6))) * t * t / r);
}

// ---------------------------------------------------------------- grid hash
// f32 -> i32 conversion of WGSL: truncation, clamped to the i32 range; NaN -> 0
inline int32_t f2i(float f) {
  if (!(f == f)) return 0;
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f <= -2147483648.0f) return INT32_MIN;
  return (int32_t)f;
}
// f32 -> u32: truncation, clamped to [0, 2^32-1]; NaN -> 0
inline u32 f2u(float f) {
  if (!(f == f) || f <= 0.0f) return 0u;
  if (f >= 4294967296.0f) return 0xffffffffu;
  return (u32)f;
}
struct i3 { int32_t x, y, z; };
// common.js:97-100  (evaluated in f32 for both instantiations)
template <class T> inline i3 cellCoord(const Prm<T>& P, float px, float py, float pz) {
  float p[3] = {px, py, pz};
  int32_t c[3];
  for (int k = 0; k < 3; k++) {
    float q = std::floor((p[k] - P.fOrigin[k]) / P.fCell);
    int32_t v = f2i(q);
    c[k] = std::min(std::max(v, 0), P.dims[k] - 1);
  }
  return i3{c[0], c[1], c[2]};
}
// common.js:101-103
template <class T> inline u32 cellIndex(const Prm<T>& P, i3 c) {
  return u32(c.x) + u32(c.y) * u32(P.dims[0]) + u32(c.z) * u32(P.dims[0] * P.dims[1]);
}

// ---------------------------------------------------------------- svd3
template <class T> struct SVD { m3<T> U; v3<T> sig; m3<T> V; };

// common.js:108-143
template <class T> inline void jacobiRot(m3<T>& Sm, m3<T>& Vm, int p, int q) {
  T Spq = Sm.c[q][p];
  T Spp = Sm.c[p][p];
  T Sqq = Sm.c[q][q];
  if (std::fabs(Spq) < T(1e-12)) return;
  T tau = (Sqq - Spp) / (T(2) * Spq);
  T t;
  if (tau >= T(0)) t = T(1) / (tau + std::sqrt(T(1) + tau * tau));
  else t = T(1) / (tau - std::sqrt(T(1) + tau * tau));
  T c = T(1) / std::sqrt(T(1) + t * t);
  T s = t * c;
  m3<T> Sn = Sm;
  for (int k = 0; k < 3; k++) {
    T Skp = Sm.c[p][k];
    T Skq = Sm.c[q][k];
    Sn.c[p][k] = c * Skp - s * Skq;
    Sn.c[q][k] = s * Skp + c * Skq;
  }
  m3<T> S2 = Sn;
  for (int k = 0; k < 3; k++) {
    T Spk = Sn.c[k][p];
    T Sqk = Sn.c[k][q];
    S2.c[k][p] = c * Spk - s * Sqk;
    S2.c[k][q] = s * Spk + c * Sqk;
  }
  Sm = S2;
  m3<T> Vn = Vm;
  for (int k = 0; k < 3; k++) {
    T Vkp = Vm.c[p][k];
    T Vkq = Vm.c[q][k];
    Vn.c[p][k] = c * Vkp - s * Vkq;
    Vn.c[q][k] = s * Vkp + c * Vkq;
  }
  Vm = Vn;
}

// common.js:145-180
template <class T> inline SVD<T> svd3(const m3<T>& F) {
  m3<T> S = mul(transpose(F), F);
  m3<T> Vm = mident<T>();
  for (int sweep = 0; sweep < 6; sweep++) {
    jacobiRot(S, Vm, 0, 1);
    jacobiRot(S, Vm, 0, 2);
    jacobiRot(S, Vm, 1, 2);
  }
  v3<T> sig = V(std::sqrt(wmax(S.c[0][0], T(0))), std::sqrt(wmax(S.c[1][1], T(0))),
                std::sqrt(wmax(S.c[2][2], T(0))));
  v3<T> v0 = Vm.c[0], v1 = Vm.c[1], v2 = Vm.c[2];
  if (sig.x < sig.y) { std::swap(sig.x, sig.y); std::swap(v0, v1); }
  if (sig.x < sig.z) { std::swap(sig.x, sig.z); std::swap(v0, v2); }
  if (sig.y < sig.z) { std::swap(sig.y, sig.z); std::swap(v1, v2); }
  if (dot(cross(v0, v1), v2) < T(0)) v2 = -v2;
  m3<T> Vout{{v0, v1, v2}};
  v3<T> u0, u1, u2;
  const T eps = T(1e-8);
  if (sig.x > eps) u0 = normalize(mul(F, v0));
  else u0 = V<T>(1, 0, 0);
  if (sig.y > eps) {
    v3<T> Fv1 = mul(F, v1);
    u1 = Fv1 - dot(Fv1, u0) * u0;
    T l = length(u1);
    if (l > eps) u1 = u1 / l;
    else u1 = normalize(cross(V<T>(0, 1, 0), u0) + splat(T(1e-4)));
  } else {
    v3<T> a = cross(u0, V<T>(0, 1, 0));
    if (length(a) < T(1e-4)) a = cross(u0, V<T>(1, 0, 0));
    u1 = normalize(a);
  }
  u2 = cross(u0, u1);
  if (det(F) < T(0)) sig.z = -sig.z;
  return SVD<T>{m3<T>{{u0, u1, u2}}, sig, Vout};
}

// common.js:182-184  U diag(d) V^T
template <class T> inline m3<T> diagMul(const m3<T>& U, v3<T> d, const m3<T>& Vm) {
  m3<T> A{{U.c[0] * d.x, U.c[1] * d.y, U.c[2] * d.z}};
  return mul(A, transpose(Vm));
}

// ---------------------------------------------------------------- StVK-Hencky
// common.js:187-189
template <class T> inline v3<T> safeSig(v3<T> s) { return vclamp(s, T(1e-3), T(1e3)); }
// common.js:190-194
template <class T> inline T stvkPsi(v3<T> sig, T mu, T lam) {
  v3<T> e = vlog(safeSig(sig));
  T tr = e.x + e.y + e.z;
  return mu * dot(e, e) + T(0.5) * lam * tr * tr;
}
// common.js:196-201
template <class T> inline v3<T> stvkDPsi(v3<T> sig, T mu, T lam) {
  v3<T> s = safeSig(sig);
  v3<T> e = vlog(s);
  T tr = e.x + e.y + e.z;
  return (T(2) * mu * e + splat(lam * tr)) / s;
}

// ---------------------------------------------------------------- return maps
// common.js:206-221
template <class T> inline v3<T> rmDruckerPrager(v3<T> sig, const Mat<T>& m) {
  T alpha = m.p0, coh = m.p1;
  v3<T> e = vlog(safeSig(sig)) - splat(coh);
  T tr = e.x + e.y + e.z;
  v3<T> dev = e - splat(tr / T(3));
  T dn = length(dev);
  if (dn < T(1e-9) || tr > T(0)) return vexp(splat(coh));
  T dg = dn + alpha * (T(3) * m.lam + T(2) * m.mu) / (T(2) * m.mu) * tr;
  if (dg <= T(0)) return sig;
  v3<T> en = e - (dg / dn) * dev;
  return vexp(en + splat(coh));
}
// common.js:224-234
template <class T> inline v3<T> rmVonMises(v3<T> sig, const Mat<T>& m) {
  T sigY = m.p0;
  v3<T> e = vlog(safeSig(sig));
  T tr = e.x + e.y + e.z;
  v3<T> dev = e - splat(tr / T(3));
  T dn = length(dev);
  T dg = dn - sigY / (T(2) * m.mu);
  if (dg <= T(0) || dn < T(1e-9)) return sig;
  v3<T> en = e - (dg / dn) * dev;
  return vexp(en);
}
// common.js:244-280.
// hint in [0, 1] (learned warm start, no reference counterpart): a predicted
// root position t in s = sY + t (sPre - sY); the bisection starts from
// [t - 1/64, t + 1/64] when g changes sign across it (13 halvings), else from
// the full bracket (18, common.js:264-265).  *tOut: the root position found
// (NaN when no bisection ran).
template <class T> inline v3<T> rmHerschelBulkley(v3<T> sig, const Mat<T>& m, T dt, T hint = T(-1),
                                                  T* tOut = nullptr) {
  T sigY = m.p0, hexp = m.p1, eta = m.p2;
  if (tOut) *tOut = std::numeric_limits<T>::quiet_NaN();
  v3<T> s3 = safeSig(sig);
  T J = s3.x * s3.y * s3.z;
  v3<T> bbar = s3 * s3 * std::pow(J, T(-2.0 / 3.0));
  T trb = bbar.x + bbar.y + bbar.z;
  v3<T> devb = bbar - splat(trb / T(3));
  T sPre = m.mu * length(devb);
  T sY = T(std::sqrt(2.0 / 3.0)) * sigY;
  if (sPre <= sY || sPre < T(1e-12)) return sig;
  T muT = (trb / T(3)) * m.mu;
  T sNew;
  if (std::fabs(hexp - T(1)) < T(1e-3) || eta < T(1e-9)) {
    sNew = sPre - (sPre - sY) / (T(1) + eta / (T(2) * muT * dt));
  } else {
    T ih = T(1) / hexp;
    T ei = std::pow(eta, ih);
    auto g = [&](T x) { return ei * (x - sPre) + T(2) * muT * dt * std::pow(wmax(x - sY, T(0)), ih); };
    T lo = sY, hi = sPre;
    int steps = 18;
    if (hint >= T(0) && hint <= T(1)) {
      const T a = wmax(hint - T(1.0 / 64.0), T(0)), b = wmin(hint + T(1.0 / 64.0), T(1));
      const T lo2 = sY + a * (sPre - sY), hi2 = sY + b * (sPre - sY);
      if ((a == T(0) || !(g(lo2) > T(0))) && (b == T(1) || g(hi2) > T(0))) {
        lo = lo2;
        hi = hi2;
        steps = 13;
      }
    }
    for (int it = 0; it < steps; it++) {
      T mid = T(0.5) * (lo + hi);
      if (g(mid) > T(0)) hi = mid;
      else lo = mid;
    }
    sNew = T(0.5) * (lo + hi);
    if (tOut) *tOut = (sNew - sY) / (sPre - sY);
  }
  v3<T> bNew = devb * (sNew / sPre) + splat(trb / T(3));
  bNew = vmax(bNew, T(1e-6));
  T detB = bNew.x * bNew.y * bNew.z;
  bNew = bNew * std::pow(detB, T(-1.0 / 3.0));
  return vsqrt(bNew) * std::pow(J, T(1.0 / 3.0));
}
// common.js:283-285
template <class T> inline v3<T> rmSnowClamp(v3<T> sig, const Mat<T>& m) {
  return {wclamp(sig.x, T(1) - m.p0, T(1) + m.p1), wclamp(sig.y, T(1) - m.p0, T(1) + m.p1),
          wclamp(sig.z, T(1) - m.p0, T(1) + m.p1)};
}
// common.js:293-364
template <class T> inline v3<T> rmNACC(v3<T> sig, const Mat<T>& m, T alphaIn, T* alphaOut) {
  T beta = m.p1, xi = m.p2, Mf = m.p3;
  T mu = m.mu;
  T kappa = (T(2.0 / 3.0)) * mu + m.lam;
  T alpha = alphaIn;
  v3<T> s3 = safeSig(sig);
  T Jtr = s3.x * s3.y * s3.z;
  v3<T> b2 = s3 * s3;
  T trb = b2.x + b2.y + b2.z;
  v3<T> devb = b2 - splat(trb / T(3));
  v3<T> sHat = mu * std::pow(Jtr, T(-2.0 / 3.0)) * devb;
  T sNorm = length(sHat);
  T q = T(std::sqrt(1.5)) * sNorm;
  T p = T(-0.5) * kappa * (Jtr * Jtr - T(1));
  alpha = wclamp(alpha, T(-0.25), T(0.4));
  T p0 = kappa * (T(0.00001) + std::sinh(wmin(xi * wmax(-alpha, T(0)), T(4))));
  T b = T(1) + T(2) * beta;
  v3<T> outSig = sig;
  if (p > p0) {
    T Jn = std::sqrt(wmax(T(-2) * p0 / kappa + T(1), T(1e-6)));
    outSig = splat(std::pow(Jn, T(1.0 / 3.0)));
    alpha = alpha + std::log(Jtr / Jn);
  } else if (p < -beta * p0) {
    T Jn = std::sqrt(wmax(T(2) * beta * p0 / kappa + T(1), T(1e-6)));
    outSig = splat(std::pow(Jn, T(1.0 / 3.0)));
    alpha = alpha + std::log(Jtr / Jn);
  } else {
    T y = b * q * q + Mf * Mf * (p + beta * p0) * (p - p0);
    if (y > T(1e-4) && sNorm > T(1e-9)) {
      if (p0 > T(1e-4) && p < p0 - T(1e-4) && p > -beta * p0 + T(1e-4)) {
        T pc = (T(1) - beta) * p0 * T(0.5);
        T dpc = pc - p, dq = -q;
        T lenDir = std::sqrt(dpc * dpc + dq * dq);
        if (lenDir > T(1e-9)) {
          T Dp = (pc - p) / lenDir;
          T Dq = -q / lenDir;
          T Aq = Mf * Mf * Dp * Dp + b * Dq * Dq;
          T Bq = Mf * Mf * Dp * (T(2) * pc - p0 + beta * p0);
          T Cq = Mf * Mf * (pc + beta * p0) * (pc - p0);
          T disc = Bq * Bq - T(4) * Aq * Cq;
          if (disc > T(0) && std::fabs(Aq) > T(1e-12)) {
            T l1 = (-Bq + std::sqrt(disc)) / (T(2) * Aq);
            T l2 = (-Bq - std::sqrt(disc)) / (T(2) * Aq);
            T px1 = pc + l1 * Dp;
            T px2 = pc + l2 * Dp;
            T px = px1;
            if ((p - pc) * (px2 - pc) > T(0)) px = px2;
            T Jx2 = T(-2) * px / kappa + T(1);
            if (Jx2 > T(1e-8)) {
              T Jx = std::sqrt(Jx2);
              if (Jx > T(1e-4)) alpha = alpha + std::log(Jtr / Jx);
            }
          }
        }
      }
      T qN = Mf * std::sqrt(wmax(-(p + beta * p0) * (p - p0), T(0)) / b);
      T sNewNorm = qN / T(std::sqrt(1.5));
      v3<T> bNew = (std::pow(Jtr, T(2.0 / 3.0)) / mu) * (sNewNorm / sNorm) * sHat + splat(trb / T(3));
      bNew = vmax(bNew, T(1e-6));
      outSig = vsqrt(bNew);
    }
  }
  *alphaOut = alpha;
  return outSig;
}
// common.js:367-375
template <class T> inline Mat<T> hardenedMat(Mat<T> m, T state) {
  if (m.kind == MAT_SNOW) {
    T s = std::exp(wmin(m.p2 * (T(1) - state), T(5)));
    m.mu = m.mu * s;
    m.lam = m.lam * s;
  }
  return m;
}
// common.js:379-396
template <class T> inline v3<T> returnMap(v3<T> sig, const Mat<T>& m, T state, T dt, T* stateOut,
                                          T hbHint = T(-1), T* hbT = nullptr) {
  switch (m.kind) {
    case 1: *stateOut = state; return rmDruckerPrager(sig, m);
    case 2: return rmNACC(sig, m, state, stateOut);
    case 3: *stateOut = state; return rmVonMises(sig, m);
    case 4: *stateOut = state; return rmHerschelBulkley(sig, m, dt, hbHint, hbT);
    case 5: {
      Mat<T> mm = hardenedMat(m, state);
      v3<T> sNew = rmSnowClamp(sig, mm);
      T jOld = sig.x * sig.y * sig.z;
      T jNew = sNew.x * sNew.y * sNew.z;
      *stateOut = wclamp(state * jOld / wmax(jNew, T(1e-6)), T(0.2), T(2.0));
      return sNew;
    }
    default: *stateOut = state; return sig;
  }
}

// ---------------------------------------------------------------- state
// Reference buffer layout (sim.js:32-50): pos/vel/aux are vec4 per particle,
// F/L/Sp are 3 vec4-padded columns (12 scalars) per particle.
template <class T> struct State {
  u32 N = 0, C = 0;
  T* pos = nullptr;   // 4N
  T* velA = nullptr;  // 4N
  T* velB = nullptr;  // 4N
  T* F = nullptr;     // 12N
  T* L = nullptr;     // 12N
  T* aux = nullptr;   // 4N
  T* Sp = nullptr;    // 12N
  u32* cellTab = nullptr;    // C
  u32* sortedIdx = nullptr;  // N
  u32* id = nullptr;         // N (persistent particle id, permuted along)
};

template <class T> inline m3<T> loadM(const T* a, u32 i) {
  const T* p = a + 12 * (size_t)i;
  return m3<T>{{V(p[0], p[1], p[2]), V(p[4], p[5], p[6]), V(p[8], p[9], p[10])}};
}
template <class T> inline void storeM(T* a, u32 i, const m3<T>& m) {
  T* p = a + 12 * (size_t)i;
  for (int j = 0; j < 3; j++) { p[4 * j] = m.c[j].x; p[4 * j + 1] = m.c[j].y; p[4 * j + 2] = m.c[j].z; p[4 * j + 3] = T(0); }
}
template <class T> inline v3<T> xyz(const T* a, u32 i) {
  const T* p = a + 4 * (size_t)i;
  return V(p[0], p[1], p[2]);
}

// ---- warm-start MLP forward (no reference counterpart), the model of
// oracle/mlp_oracle.py: x = bf16(features), h1 = bf16(relu(W1 x + b1)),
// h2 = bf16(relu(W2 h1 + b2)), y = W3 h2 + b3; weights bf16-rounded from the
// canonical f32 vector W1[64][12], b1, W2[64][64], b2, W3[2][64], b3; sums in
// double, rounded to float where the kernels hold f32 values.
inline float bf16r(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return f;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}
template <class T> inline void mlpForward(const float* p, const T* x12, double y[2]) {
  float x[12], h1[64], h2[64];
  for (int f = 0; f < 12; f++) x[f] = bf16r(float(x12[f]));
  for (int u = 0; u < 64; u++) {
    double z = 0;
    for (int f = 0; f < 12; f++) z += double(bf16r(p[u * 12 + f])) * double(x[f]);
    float zf = float(z + double(p[768 + u]));
    h1[u] = bf16r(zf > 0.0f ? zf : 0.0f);
  }
  for (int o = 0; o < 64; o++) {
    double z = 0;
    for (int i = 0; i < 64; i++) z += double(bf16r(p[832 + o * 64 + i])) * double(h1[i]);
    float zf = float(z + double(p[4928 + o]));
    h2[o] = bf16r(zf > 0.0f ? zf : 0.0f);
  }
  for (int k = 0; k < 2; k++) {
    double z = 0;
    for (int u = 0; u < 64; u++) z += double(bf16r(p[4992 + k * 64 + u])) * double(h2[u]);
    y[k] = double(float(z + double(p[5120 + k])));
  }
}

// Stage names for partial substeps (used by per-kernel parity tests)
enum Stage { ST_SORT = 1, ST_REORDER = 2, ST_COMPUTE_L = 3, ST_SOLVE = 4, ST_XSPH = 5, ST_UPDATE_F = 6, ST_FULL = 7 };

template <class T> struct Sim {
  Prm<T> P;
  Mat<T> M[8];
  State<T> S;
  int nthreads = 1;
  bool flip = false;  // false: velA is "in"
  // optional per-iteration residual output (sum over solids of numerator^2, C^2)
  double* resid = nullptr;  // 2 entries per iteration
  // residual-driven loop (tol > 0: `iterations` is a cap; no reference counterpart)
  double tol = 0;
  int itersDone = 0;
  // learned warm start (no reference counterpart): 0 plain, 1 learned, 2 collect
  // pairs, 3 a given lambda0 per particle id (the warm-start ceiling experiments)
  int wsMode = 0;
  const float* wsParams = nullptr;  // canonical MLP parameters (mode 1)
  T* feat = nullptr;                // N x 16 feature rows (modes 1, 2, 3)
  T* target = nullptr;              // N x 2 targets (mode 2)
  const T* lam0ById = nullptr;      // lambda0 indexed by persistent particle id (mode 3)
  std::vector<T> lamPrev, hbHint;
  T* velIn() { return flip ? S.velB : S.velA; }
  T* velOut() { return flip ? S.velA : S.velB; }

  u32 matOf(u32 i) const { return f2u(float(S.pos[4 * (size_t)i + 3])); }
  bool isSolid(u32 mi) const { return M[mi & 7].kind != MAT_WATER; }

  // simkernels.js:62-74
  void rowRange(i3 c, u32& beg, u32& end) const {
    if (c.y < 0 || c.y >= P.dims[1] || c.z < 0 || c.z >= P.dims[2]) { beg = end = 0; return; }
    int32_t x0 = std::max(c.x - 1, 0);
    int32_t x1 = std::min(c.x + 1, P.dims[0] - 1);
    u32 c0 = cellIndex(P, i3{x0, c.y, c.z});
    u32 c2 = cellIndex(P, i3{x1, c.y, c.z});
    beg = S.cellTab[c0];
    u32 e = P.N;
    if (c2 + 1u < P.numCells) e = S.cellTab[c2 + 1u];
    end = std::min(e, beg + 384u);
  }
  i3 ccOf(v3<T> x) const { return cellCoord(P, float(x.x), float(x.y), float(x.z)); }
  template <class Fn> void forRows(v3<T> xi, Fn fn) const {
    i3 cc = ccOf(xi);
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++) {
        u32 b0, b1;
        rowRange(i3{cc.x, cc.y + dy, cc.z + dz}, b0, b1);
        for (u32 b = b0; b < b1; b++) fn(b);
      }
  }

  // ---- counting sort: simkernels.js:78-147 (stable within-cell order) + the
  // per-particle work of scatter (Vn, predictor, lambda reset)
  void sortAndScatter() {
    const u32 N = P.N, C = P.numCells;
    std::vector<u32> key(N);
    std::vector<u32> cnt(C, 0u);
    for (u32 i = 0; i < N; i++) {
      key[i] = cellIndex(P, cellCoord(P, float(S.pos[4 * i]), float(S.pos[4 * i + 1]), float(S.pos[4 * i + 2])));
      cnt[key[i]]++;
    }
    // exclusive prefix (what cellTab holds after the atomicSub scatter)
    u32 run = 0;
    for (u32 c = 0; c < C; c++) { S.cellTab[c] = run; run += cnt[c]; }
    std::vector<u32> cursor(S.cellTab, S.cellTab + C);
    for (u32 i = 0; i < N; i++) S.sortedIdx[cursor[key[i]]++] = i;
    T* vin = velIn();
    if (wsMode) {  // lambda at the end of the previous substep, before its reset
      lamPrev.resize(N);
      for (u32 i = 0; i < N; i++) lamPrev[i] = S.aux[4 * (size_t)i + 1];
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long ii = 0; ii < (long)N; ii++) {
      u32 i = (u32)ii;
      T Vn = P.V0 * wclamp(det(loadM(S.F, i)), T(0.1), T(10.0));
      S.aux[4 * i] = Vn;
      v3<T> v = xyz(vin, i) + P.dt * P.gravity;
      vin[4 * i] = v.x; vin[4 * i + 1] = v.y; vin[4 * i + 2] = v.z; vin[4 * i + 3] = Vn;
      S.aux[4 * i + 1] = T(0);
    }
  }
  // simkernels.js:154-186 (staging copies), plus the id channel
  void reorder() {
    const u32 N = P.N;
    std::vector<T> pos(4 * (size_t)N), vel(4 * (size_t)N), aux(4 * (size_t)N), F(12 * (size_t)N);
    std::vector<u32> id(N);
    T* vin = velIn();
    std::vector<T> lp(wsMode ? N : 0);
    for (u32 j = 0; j < N; j++) {
      u32 i = S.sortedIdx[j];
      if (wsMode) lp[j] = lamPrev[i];
      std::memcpy(&pos[4 * (size_t)j], S.pos + 4 * (size_t)i, 4 * sizeof(T));
      std::memcpy(&vel[4 * (size_t)j], vin + 4 * (size_t)i, 4 * sizeof(T));
      std::memcpy(&aux[4 * (size_t)j], S.aux + 4 * (size_t)i, 4 * sizeof(T));
      std::memcpy(&F[12 * (size_t)j], S.F + 12 * (size_t)i, 12 * sizeof(T));
      id[j] = S.id[i];
    }
    std::memcpy(S.pos, pos.data(), pos.size() * sizeof(T));
    std::memcpy(vin, vel.data(), vel.size() * sizeof(T));
    std::memcpy(velOut(), vel.data(), vel.size() * sizeof(T));
    std::memcpy(S.aux, aux.data(), aux.size() * sizeof(T));
    for (u32 j = 0; j < N; j++)  // F columns; pad lanes are written 0 by stageF
      for (int c = 0; c < 3; c++) {
        for (int r = 0; r < 3; r++) S.F[12 * (size_t)j + 4 * c + r] = F[12 * (size_t)j + 4 * c + r];
        S.F[12 * (size_t)j + 4 * c + 3] = T(0);
      }
    std::memcpy(S.id, id.data(), N * sizeof(u32));
    if (wsMode) lamPrev.swap(lp);
  }
  // simkernels.js:189-221
  void computeL() {
    const T* vin = velIn();
    nbrCnt.assign(P.N, 0u);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (long ii = 0; ii < (long)P.N; ii++) {
      u32 i = (u32)ii;
      v3<T> xi = xyz(S.pos, i);
      m3<T> A = mzero<T>();
      forRows(xi, [&](u32 b) {
        if (b != i) {
          v3<T> dxv = xi - xyz(S.pos, b);
          if (dot(dxv, dxv) < P.support * P.support) {
            nbrCnt[i]++;
            v3<T> g = kernelGradW(P, dxv);
            v3<T> r = xyz(S.pos, b) - xi;
            A = madd(A, mscale(outer(g, r), vin[4 * (size_t)b + 3]));
          }
        }
      });
      SVD<T> s = svd3(A);
      const T tol = T(0.2);
      v3<T> inv = splat<T>(0);
      if (std::fabs(s.sig.x) > tol) inv.x = T(1) / s.sig.x;
      if (std::fabs(s.sig.y) > tol) inv.y = T(1) / s.sig.y;
      if (std::fabs(s.sig.z) > tol) inv.z = T(1) / s.sig.z;
      storeM(S.L, i, diagMul(s.V, inv, s.U));
    }
  }
  // simkernels.js:224-245
  m3<T> velocityGradient(u32 i, v3<T> xi) const {
    m3<T> G = mzero<T>();
    m3<T> Li = loadM(S.L, i);
    const T* vin = flip ? S.velB : S.velA;
    v3<T> vi = xyz(vin, i);
    forRows(xi, [&](u32 b) {
      if (b != i && isSolid(matOf(b))) {
        v3<T> dxv = xi - xyz(S.pos, b);
        if (dot(dxv, dxv) < P.support * P.support) {
          v3<T> gt = mul(Li, kernelGradW(P, dxv));
          v3<T> dv = (xyz(vin, b) - vi) * vin[4 * (size_t)b + 3];
          G = madd(G, outer(dv, gt));
        }
      }
    });
    return G;
  }
  void storeSpSelf(u32 i, const m3<T>& m, v3<T> s) {
    T* p = S.Sp + 12 * (size_t)i;
    for (int j = 0; j < 3; j++) { p[4 * j] = m.c[j].x; p[4 * j + 1] = m.c[j].y; p[4 * j + 2] = m.c[j].z; p[4 * j + 3] = s[j]; }
  }
  // HB bracket hint of particle i (learned warm start), -1 for none
  T hintOf(u32 i) const {
    return (wsMode == 1 && !hbHint.empty() && M[matOf(i) & 7].kind == MAT_HB) ? hbHint[i] : T(-1);
  }
  // simkernels.js:255-300 for solid particle i at the current velocity: trial F,
  // SVD, return map, C and, when C >= 1e-7, GL and the second sweep's denom,
  // sgrad and cdot
  struct Cons {
    bool active;
    T C, denom, cdot, hbT;
    m3<T> GL;
    v3<T> sgrad, sigTrial, sigPost;
  };
  Cons constraintAt(u32 i, T hint, bool wantHbT) const {
    Cons c;
    c.active = false;
    const T* vin = flip ? S.velB : S.velA;
    u32 mi = matOf(i);
    Mat<T> m = hardenedMat(M[mi & 7], S.aux[4 * (size_t)i + 2]);
    v3<T> xi = xyz(S.pos, i);
    m3<T> Fn = loadM(S.F, i);
    m3<T> gradV = velocityGradient(i, xi);
    m3<T> Ftr = mul(madd(mident<T>(), mscale(gradV, P.dt)), Fn);
    SVD<T> sv = svd3(Ftr);
    T st;
    c.hbT = std::numeric_limits<T>::quiet_NaN();
    v3<T> sig = returnMap(sv.sig, m, S.aux[4 * (size_t)i + 2], P.dt, &st, hint, wantHbT ? &c.hbT : nullptr);
    c.sigTrial = sv.sig;
    c.sigPost = sig;
    T psi = stvkPsi(sig, m.mu, m.lam);
    c.C = std::sqrt(wmax(T(2) * psi, T(0)));
    if (c.C < T(1e-7)) return c;
    c.active = true;
    v3<T> dpsi = stvkDPsi(sig, m.mu, m.lam);
    m3<T> Pk = diagMul(sv.U, dpsi / c.C, sv.V);
    c.GL = mul(mul(Pk, transpose(Fn)), loadM(S.L, i));
    T denom = 0, cdot = 0;
    v3<T> sgrad = splat<T>(0);
    forRows(xi, [&](u32 b) {
      if (b != i && isSolid(matOf(b))) {
        v3<T> dxv = xi - xyz(S.pos, b);
        if (dot(dxv, dxv) < P.support * P.support) {
          v3<T> vb = xyz(vin, b);
          v3<T> gb = vin[4 * (size_t)b + 3] * mul(c.GL, kernelGradW(P, dxv));
          denom = denom + dot(gb, gb) / M[matOf(b) & 7].mass;
          sgrad = sgrad - gb;
          cdot = cdot + dot(gb, vb);
        }
      }
    });
    c.denom = denom + dot(sgrad, sgrad) / m.mass;
    c.cdot = cdot + dot(sgrad, xyz(vin, i));
    c.sgrad = sgrad;
    return c;
  }
  T alphaT() const { return (T(1) / P.V0) / (P.dt * P.dt); }
  // simkernels.js:248-309  (+ residual numerator for the convergence monitor).
  // dlamOverride (the warm start's lambda0) replaces the Jacobi step and the
  // HB hint is not used.
  double lastR2 = 0, lastC2 = 0;
  void solveA(int iter, const T* dlamOverride = nullptr) {
    double r2 = 0, c2 = 0;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256) reduction(+ : r2, c2)
    for (long ii = 0; ii < (long)P.N; ii++) {
      u32 i = (u32)ii;
      if (!isSolid(matOf(i))) { storeSpSelf(i, mzero<T>(), splat<T>(0)); continue; }
      const Cons c = constraintAt(i, dlamOverride ? T(-1) : hintOf(i), false);
      if (!c.active) { storeSpSelf(i, mzero<T>(), splat<T>(0)); continue; }
      T g = P.kDamp;
      T lam = S.aux[4 * (size_t)i + 1];
      T num = -c.C - alphaT() * lam - g * P.dt * c.cdot;
      T dlam = dlamOverride ? dlamOverride[i] : P.omega * num / ((T(1) + g) * c.denom + alphaT());
      S.aux[4 * (size_t)i + 1] = lam + dlam;
      storeSpSelf(i, mscale(c.GL, dlam), c.sgrad * dlam);
      r2 += double(num) * double(num);
      c2 += double(c.C) * double(c.C);
    }
    lastR2 = r2;
    lastC2 = c2;
    if (resid && iter >= 0) { resid[2 * iter] = r2; resid[2 * iter + 1] = c2; }
  }
  // ---- learned warm start (no reference counterpart)
  // feature rows at lambda = 0; returns sum C^2 over active solids
  std::vector<u32> nbrCnt;  // support neighbours per particle (computeL)
  double features() {
    double c2 = 0;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256) reduction(+ : c2)
    for (long ii = 0; ii < (long)P.N; ii++) {
      u32 i = (u32)ii;
      T* row = feat + 16 * (size_t)i;
      for (int k = 0; k < 16; k++) row[k] = T(0);
      row[13] = std::numeric_limits<T>::quiet_NaN();
      u32 mi = matOf(i);
      if (!isSolid(mi)) continue;
      const Cons c = constraintAt(i, T(-1), true);
      if (!c.active) continue;
      Mat<T> m = hardenedMat(M[mi & 7], S.aux[4 * (size_t)i + 2]);
      T g = P.kDamp;
      T den = (T(1) + g) * c.denom + alphaT();
      T d0 = (-c.C - g * P.dt * c.cdot) / den;
      v3<T> lt = vlog(safeSig(c.sigTrial)), lp = vlog(safeSig(c.sigPost));
      T trp = lp.x + lp.y + lp.z;
      T strength = 0;
      switch (m.kind) {
        case MAT_DP: strength = m.p0; break;
        case MAT_VM:
        case MAT_HB: strength = T(100) * m.p0 / m.mu; break;
        case MAT_NACC: strength = T(0.1) * m.p2; break;
        case MAT_SNOW: strength = T(10) * m.p0; break;
        default: break;
      }
      row[0] = T(100) * lt.x; row[1] = T(100) * lt.y; row[2] = T(100) * lt.z;
      row[3] = T(100) * trp;
      row[4] = T(100) * length(lp - splat(trp / T(3)));
      row[5] = std::log10(m.mu) - T(5);
      row[6] = m.lam / (m.lam + T(2) * m.mu);
      row[7] = T(m.kind) / T(6);
      row[8] = strength;
      row[9] = T(nbrCnt[i]) / T(40);
      row[10] = (d0 != T(0)) ? wclamp(lamPrev[i] / d0, T(-8), T(8)) : T(0);
      row[11] = alphaT() / den;
      row[12] = d0;
      row[13] = c.hbT;
      row[14] = c.C;
      row[15] = -c.C - g * P.dt * c.cdot;
      c2 += double(c.C) * double(c.C);
    }
    return c2;
  }
  // simkernels.js:369-404
  v3<T> applyBoundary(v3<T> xcIn, v3<T> vIn, v3<T> xn, T fric) const {
    v3<T> vv = vIn, xq = xcIn;
    for (int k = 0; k < 3; k++) {
      T lo = P.domainMin[k], hi = P.domainMax[k];
      if (xq[k] < lo || xq[k] > hi) {
        T bound = (xq[k] < lo) ? lo : hi;
        T vnNew = (bound - xn[k]) / P.dt;
        T dvn = std::fabs(vnNew - vv[k]);
        v3<T> vt = vv;
        vt[k] = T(0);
        T vtl = length(vt);
        if (vtl > T(1e-9)) {
          T drop = wmin(fric * dvn, vtl);
          vv = vv - vt / vtl * drop;
        }
        vv[k] = vnNew;
        xq = xn + P.dt * vv;
      }
    }
    if (P.colliderActive > T(0.5)) {
      v3<T> d = xq - P.colliderPos;
      T dist = length(d);
      if (dist < P.colliderR && dist > T(1e-9)) {
        v3<T> n = d / dist;
        T vOut = dot(vv - P.colliderVel, n);
        T vGoal = wmin((P.colliderR - dist) / P.dt, T(0.6));
        if (vOut < vGoal) vv = vv + n * (vGoal - vOut);
      }
    }
    return vv;
  }
  // simkernels.js:315-366
  void solveB() {
    const T* vin = velIn();
    T* vout = velOut();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (long ii = 0; ii < (long)P.N; ii++) {
      u32 i = (u32)ii;
      u32 mi = matOf(i);
      v3<T> dv = splat<T>(0), corr = splat<T>(0);
      v3<T> xi = xyz(S.pos, i);
      const Mat<T>& m = M[mi & 7];
      bool solid = isSolid(mi);
      if (solid) {
        const T* sp = S.Sp + 12 * (size_t)i;
        dv = V(sp[3], sp[7], sp[11]);
        T Vi = vin[4 * (size_t)i + 3];
        v3<T> xci = xi + P.dt * xyz(vin, i);
        T minD = T(0.75) * P.h;
        forRows(xi, [&](u32 b) {
          if (b != i && isSolid(matOf(b))) {
            v3<T> pb = xyz(S.pos, b);
            v3<T> dxp = pb - xi;
            if (dot(dxp, dxp) < P.support * P.support) {
              dv = dv + Vi * mul(loadM(S.Sp, b), kernelGradW(P, dxp));
              v3<T> xcb = pb + P.dt * xyz(vin, b);
              v3<T> dxc = xci - xcb;
              T dist = length(dxc);
              if (dist < minD && dist > T(1e-9)) {
                T mb = M[matOf(b) & 7].mass;
                T w = mb / (m.mass + mb);
                corr = corr + dxc / dist * ((minD - dist) * w);
              }
            }
          }
        });
      }
      v3<T> dvv = dv / (m.mass * P.dt);
      T dvl = length(dvv);
      T dvCap = T(0.2) * P.support / P.dt;
      if (dvl > dvCap) dvv = dvv * (dvCap / dvl);
      v3<T> v = xyz(vin, i) + dvv + corr * (P.contactOmega / P.dt);
      v = applyBoundary(xi + P.dt * v, v, xi, m.boundFriction);
      vout[4 * (size_t)i] = v.x; vout[4 * (size_t)i + 1] = v.y; vout[4 * (size_t)i + 2] = v.z;
      vout[4 * (size_t)i + 3] = vin[4 * (size_t)i + 3];
    }
    flip = !flip;
  }
  // simkernels.js:407-439
  void pbfA() {
    const T* vin = velIn();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (long ii = 0; ii < (long)P.N; ii++) {
      u32 i = (u32)ii;
      if (isSolid(matOf(i))) { S.aux[4 * (size_t)i + 3] = T(0); continue; }
      const Mat<T>& m = M[matOf(i) & 7];
      T rho0 = m.density;
      v3<T> xn = xyz(S.pos, i);
      v3<T> xci = xn + P.dt * xyz(vin, i);
      T rho = m.mass * kernelPoly6(P, T(0));
      v3<T> gradSum = splat<T>(0);
      T grad2 = 0;
      forRows(xn, [&](u32 b) {
        if (b != i) {
          v3<T> xcb = xyz(S.pos, b) + P.dt * xyz(vin, b);
          v3<T> dxv = xci - xcb;
          if (dot(dxv, dxv) < P.support * P.support) {
            T mb = M[matOf(b) & 7].mass;
            rho = rho + mb * kernelPoly6(P, length(dxv));
            v3<T> g = kernelSpikyGrad(P, dxv) * (mb / rho0);
            gradSum = gradSum + g;
            grad2 = grad2 + dot(g, g);
          }
        }
      });
      T C = wmax(rho / rho0 - T(1), T(0));
      S.aux[4 * (size_t)i + 3] = -C / (grad2 + dot(gradSum, gradSum) + P.pbfEps);
    }
  }
  // simkernels.js:441-477
  void pbfB() {
    const T* vin = velIn();
    T* vout = velOut();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (long ii = 0; ii < (long)P.N; ii++) {
      u32 i = (u32)ii;
      u32 mi = matOf(i);
      v3<T> xn = xyz(S.pos, i);
      v3<T> xci = xn + P.dt * xyz(vin, i);
      v3<T> dp = splat<T>(0);
      T li = S.aux[4 * (size_t)i + 3];
      bool iAmWater = !isSolid(mi);
      forRows(xn, [&](u32 b) {
        if (b != i) {
          bool bWater = !isSolid(matOf(b));
          if (iAmWater || bWater) {
            v3<T> xcb = xyz(S.pos, b) + P.dt * xyz(vin, b);
            v3<T> dxv = xci - xcb;
            if (dot(dxv, dxv) < P.support * P.support) {
              const Mat<T>& mw = M[(bWater ? matOf(b) : mi) & 7];
              T lb = S.aux[4 * (size_t)b + 3];
              T w = kernelPoly6(P, length(dxv));
              T w0 = kernelPoly6(P, T(0.2) * P.support);
              T sCorr = -P.sCorrK * std::pow(w / wmax(w0, T(1e-12)), T(4));
              dp = dp + (li + lb + sCorr) * kernelSpikyGrad(P, dxv) * (mw.mass / mw.density);
            }
          }
        }
      });
      T scale = iAmWater ? T(1) : T(0.5);
      v3<T> v = xyz(vin, i) + dp * (scale / P.dt);
      vout[4 * (size_t)i] = v.x; vout[4 * (size_t)i + 1] = v.y; vout[4 * (size_t)i + 2] = v.z;
      vout[4 * (size_t)i + 3] = vin[4 * (size_t)i + 3];
    }
    flip = !flip;
  }
  // simkernels.js:480-507
  void xsph() {
    const T* vin = velIn();
    T* vout = velOut();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (long ii = 0; ii < (long)P.N; ii++) {
      u32 i = (u32)ii;
      v3<T> xi = xyz(S.pos, i), vi = xyz(vin, i);
      bool iSolid = isSolid(matOf(i));
      v3<T> acc = splat<T>(0);
      forRows(xi, [&](u32 b) {
        if (b != i && (isSolid(matOf(b)) == iSolid)) {
          v3<T> dxv = xi - xyz(S.pos, b);
          if (dot(dxv, dxv) < P.support * P.support) {
            T Vb = vin[4 * (size_t)b + 3];
            acc = acc + Vb * (xyz(vin, b) - vi) * kernelW(P, length(dxv));
          }
        }
      });
      T c = P.xsphC;
      if (!iSolid) c = P.xsphC * T(10);
      v3<T> v = vi + c * acc;
      vout[4 * (size_t)i] = v.x; vout[4 * (size_t)i + 1] = v.y; vout[4 * (size_t)i + 2] = v.z;
      vout[4 * (size_t)i + 3] = vin[4 * (size_t)i + 3];
    }
    flip = !flip;
  }
  // simkernels.js:510-525
  void updateF() {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (long ii = 0; ii < (long)P.N; ii++) {
      u32 i = (u32)ii;
      u32 mi = matOf(i);
      if (!isSolid(mi)) continue;
      Mat<T> m = hardenedMat(M[mi & 7], S.aux[4 * (size_t)i + 2]);
      m3<T> gradV = velocityGradient(i, xyz(S.pos, i));
      m3<T> Ftr = mul(madd(mident<T>(), mscale(gradV, P.dt)), loadM(S.F, i));
      SVD<T> sv = svd3(Ftr);
      T st;
      v3<T> sig = returnMap(sv.sig, m, S.aux[4 * (size_t)i + 2], P.dt, &st, hintOf(i));
      storeM(S.F, i, diagMul(sv.U, safeSig(sig), sv.V));
      S.aux[4 * (size_t)i + 2] = st;
    }
  }
  // simkernels.js:528-542
  void integrate() {
    T* vin = velIn();
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long ii = 0; ii < (long)P.N; ii++) {
      u32 i = (u32)ii;
      v3<T> v = xyz(vin, i);
      T vmax = T(0.45) * P.support / P.dt;
      T vl = length(v);
      if (vl > vmax) v = v * (vmax / vl);
      vin[4 * (size_t)i] = v.x; vin[4 * (size_t)i + 1] = v.y; vin[4 * (size_t)i + 2] = v.z;
      v3<T> x = xyz(S.pos, i) + P.dt * v;
      x = V(wclamp(x.x, P.domainMin.x, P.domainMax.x), wclamp(x.y, P.domainMin.y, P.domainMax.y),
            wclamp(x.z, P.domainMin.z, P.domainMax.z));
      S.pos[4 * (size_t)i] = x.x; S.pos[4 * (size_t)i + 1] = x.y; S.pos[4 * (size_t)i + 2] = x.z;
    }
  }
  // sim.js:220-272 -- one substep; stops early after `stop` for staged parity
  void substep(int iterations, bool hasWater, int stop = ST_FULL) {
    flip = false;
    sortAndScatter();
    if (stop == ST_SORT) return;
    reorder();
    if (stop == ST_REORDER) return;
    computeL();
    if (stop == ST_COMPUTE_L) return;
    bool c0preset = false;
    double c0 = 0;
    int warm = 0;
    if (wsMode) {
      c0 = features();
      c0preset = true;
      if (wsMode == 1 && iterations > 0) {
        // the first Jacobi iteration with Delta-lambda = lambda0 = clamp(y0, 0, 2) d0
        // in place of omega d0, then plain Jacobi
        std::vector<T> lam0(P.N);
        hbHint.assign(P.N, T(-1));
        for (u32 i = 0; i < P.N; i++) {
          const T* row = feat + 16 * (size_t)i;
          double y[2];
          mlpForward(wsParams, row, y);
          // a NaN output gives lambda0 = 0 for the particle (plain Jacobi from zero)
          const T y0 = T(float(y[0]));
          lam0[i] = (row[12] == T(0) || y0 != y0) ? T(0) : wclamp(y0, T(0), T(2)) * row[12];
          hbHint[i] = T(float(y[1]));
        }
        solveA(-1, lam0.data());
        solveB();
        if (hasWater) { pbfA(); pbfB(); }
        warm = 1;
      }
      if (wsMode == 3 && iterations > 0) {
        // the first Jacobi iteration applies Delta-lambda = the given lambda0 of
        // each particle (zero where its constraint is inactive), then plain Jacobi
        std::vector<T> lam0(P.N);
        for (u32 i = 0; i < P.N; i++) lam0[i] = feat[16 * (size_t)i + 12] == T(0) ? T(0) : lam0ById[S.id[i]];
        solveA(-1, lam0.data());
        solveB();
        if (hasWater) { pbfA(); pbfB(); }
        warm = 1;
      }
    }
    itersDone = warm;
    for (int k = 0; k < iterations - warm; k++) {
      solveA(k);
      if (k == 0 && !c0preset) c0 = lastC2;
      const double r2 = lastR2;
      solveB();
      if (hasWater) { pbfA(); pbfB(); }
      itersDone = warm + k + 1;
      if (tol > 0 && (c0 <= 0 || r2 <= tol * tol * c0)) break;
    }
    if (wsMode == 2)
      for (u32 i = 0; i < P.N; i++) {
        const T d0 = feat[16 * (size_t)i + 12];
        target[2 * (size_t)i] = d0 != T(0) ? S.aux[4 * (size_t)i + 1] / d0 : std::numeric_limits<T>::quiet_NaN();
        target[2 * (size_t)i + 1] = feat[16 * (size_t)i + 13];
      }
    if (stop == ST_SOLVE) { finishFlip(); return; }
    xsph();
    if (stop == ST_XSPH) { finishFlip(); return; }
    updateF();
    if (stop == ST_UPDATE_F) { finishFlip(); return; }
    integrate();
    finishFlip();
  }
  // sim.js:269-271: the final velocity must end in velA
  void finishFlip() {
    if (flip) std::memcpy(S.velA, S.velB, 4 * sizeof(T) * (size_t)P.N);
    flip = false;
  }
};

}  // namespace xo
