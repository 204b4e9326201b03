// 2D fp64 XPBI engine: the algorithm of prototype/xpbi2d.py:123-241 on the GPU.
//
// The prototype's scenes are small (666 particles at its own constants, 3,200
// for the 40 x 80 column at r = 0.01), so one launch runs every substep of a call:
// a cooperative grid of 256-thread workgroups, one thread per particle, whose ~95
// phases per substep are separated by grid barriers (or, selectable, a single
// workgroup of 1024 threads with workgroup barriers).  The state (a few hundred
// KB) stays in L2 and no launch or host round trip sits between the phases.
//
// Arithmetic is float64 without contraction; every sum runs over a particle's
// neighbours in ascending index order and every expression is evaluated in the
// order the prototype's numpy code evaluates it, so that results differ from a
// float64 restatement in that order only where log/exp round differently.
//
// Neighbours: each substep re-grids x^n into cells of side 2h (counting sort
// within the workgroup), and each particle's list holds every b != p with
// |x_p - x_b|^2 <= (2h)^2 (xpbi2d.py:125-129), sorted ascending.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "xpbi.h"

#pragma clang fp contract(off)

namespace xpbi2d {

typedef uint32_t u32;
constexpr int kThreads = 1024;
constexpr u32 kMaxNbr = 48;  // the column at rest has ~10 neighbours within 2h

struct d2 {
  double x, y;
};
struct m2 {  // row-major 2x2: a = (0,0), b = (0,1), c = (1,0), d = (1,1)
  double a, b, c, d;
};

struct Prm {
  double h, support, supp2, dt, mass, V0, mu, lam, alphaDP, wsig, xsphC, omega, alphaT, mdt, dom0, dom1, g0, g1;
  double hc;  // h - 0.25 h (xpbi2d.py:193)
  int iterations, ncx, ncy;
  u32 N, C;
};

struct Buf {
  d2 *x, *v, *xc, *gCp, *dvx;
  m2 *F, *L, *G;
  double *Vn, *lam, *dlam;
  u32 *cellCount, *cellStart, *cursor, *sorted, *cellOf, *nbr, *cnt, *err;
};

// the threads that run a scene: a cooperative grid (barrier: grid sync) or one
// workgroup (barrier: workgroup barrier); phase loops stride by `size`
template <bool kGrid>
struct Team {
  u32 rank, size;
  __device__ Team()
      : rank(kGrid ? blockIdx.x * blockDim.x + threadIdx.x : threadIdx.x),
        size(kGrid ? gridDim.x * blockDim.x : blockDim.x) {}
  __device__ void sync() const {
    if (kGrid) cooperative_groups::this_grid().sync();
    else __syncthreads();
  }
};

// ---------------------------------------------------------------- math

__device__ __forceinline__ double kW(const Prm& P, double r) {  // xpbi2d.py:64-72
  const double q = r / P.h;
  const double t = 1.0 - 0.5 * q;
  const double t2 = t * t;
  return q < 2.0 ? P.wsig * (t2 * t2) * (2.0 * q + 1.0) : 0.0;
}

__device__ __forceinline__ d2 kGradW(const Prm& P, d2 dxv, double r) {  // xpbi2d.py:74-88
  const double q = r / P.h;
  const double t = 1.0 - 0.5 * q;
  const double dwdq = P.wsig * (-5.0 * q) * ((t * t) * t);
  const double coef = (q < 2.0 && r > 1e-12) ? dwdq / (P.h * r) : 0.0;
  return d2{dxv.x * coef, dxv.y * coef};
}

__device__ __forceinline__ d2 sub(d2 a, d2 b) { return d2{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ double nrm(d2 a) { return sqrt(a.x * a.x + a.y * a.y); }
__device__ __forceinline__ d2 mv(const m2& A, d2 g) { return d2{A.a * g.x + A.b * g.y, A.c * g.x + A.d * g.y}; }

// A = U diag(s0, s1) V^T: a symmetric Jacobi rotation of A^T A (Golub & Van
// Loan 8.5.2), sigma_2 = |det A| / sigma_1, det V = +1, det U = sign(det A)
struct Svd {
  m2 U, V;
  double s0, s1;
};
__device__ __forceinline__ Svd svd2(const m2& A) {
  const double m00 = A.a * A.a + A.c * A.c;
  const double m01 = A.a * A.b + A.c * A.d;
  const double m11 = A.b * A.b + A.d * A.d;
  double t = 0.0;
  if (m01 != 0.0) {
    const double tau = (m11 - m00) / (2.0 * m01);
    t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
  }
  const double c = 1.0 / sqrt(1.0 + t * t);
  const double s = t * c;
  const double e0 = m00 - t * m01;
  const double e1 = m11 + t * m01;
  Svd o;
  if (e1 > e0) o.V = m2{s, -c, c, s};
  else o.V = m2{c, s, -s, c};
  const double w0 = A.a * o.V.a + A.b * o.V.c;
  const double w1 = A.c * o.V.a + A.d * o.V.c;
  o.s0 = sqrt(w0 * w0 + w1 * w1);
  const double det = A.a * A.d - A.b * A.c;
  double u00 = 1.0, u10 = 0.0;
  o.s1 = 0.0;
  if (o.s0 > 0.0) {
    u00 = w0 / o.s0;
    u10 = w1 / o.s0;
    o.s1 = fabs(det) / o.s0;
  }
  const double sg = det < 0.0 ? -1.0 : 1.0;
  o.U = m2{u00, sg * -u10, u10, sg * u00};
  return o;
}

// U diag(s) V^T, entry (i, k) = (U_i0 s0) V_k0 + (U_i1 s1) V_k1
__device__ __forceinline__ m2 usv(const m2& U, double s0, double s1, const m2& V) {
  return m2{(U.a * s0) * V.a + (U.b * s1) * V.b, (U.a * s0) * V.c + (U.b * s1) * V.d,
            (U.c * s0) * V.a + (U.d * s1) * V.b, (U.c * s0) * V.c + (U.d * s1) * V.d};
}
__device__ __forceinline__ m2 mm(const m2& A, const m2& B) {
  return m2{A.a * B.a + A.b * B.c, A.a * B.b + A.b * B.d, A.c * B.a + A.d * B.c, A.c * B.b + A.d * B.d};
}

// Drucker-Prager return map on the singular values (xpbi2d.py:94-110)
__device__ __forceinline__ void druckerPrager(const Prm& P, double s0, double s1, double* o0, double* o1) {
  const double d = 2.0;
  const double e0 = log(fmax(s0, 1e-10)), e1 = log(fmax(s1, 1e-10));
  const double tr = e0 + e1;
  const double dv0 = e0 - tr / d, dv1 = e1 - tr / d;
  const double dn = sqrt(dv0 * dv0 + dv1 * dv1);
  const double dg = dn + ((P.alphaDP * (d * P.lam + 2.0 * P.mu)) / (2.0 * P.mu)) * tr;
  double r0 = e0, r1 = e1;
  if (tr > 0.0) {
    r0 = 0.0;
    r1 = 0.0;
  } else if (dg > 0.0 && dn > 1e-12) {
    const double scale = 1.0 - dg / dn;
    r0 = dv0 * scale + tr / d;
    r1 = dv1 * scale + tr / d;
  }
  *o0 = exp(r0);
  *o1 = exp(r1);
}

// Hencky StVK Psi and dPsi/dSigma (xpbi2d.py:112-118)
__device__ __forceinline__ double stvk(const Prm& P, double s0, double s1, double* g0, double* g1) {
  const double c0 = fmax(s0, 1e-10), c1 = fmax(s1, 1e-10);
  const double l0 = log(c0), l1 = log(c1);
  const double tr = l0 + l1;
  *g0 = ((2.0 * P.mu) * l0 + P.lam * tr) / c0;
  *g1 = ((2.0 * P.mu) * l1 + P.lam * tr) / c1;
  return P.mu * (l0 * l0 + l1 * l1) + (0.5 * P.lam) * (tr * tr);
}

// ---------------------------------------------------------------- kernel

// the neighbours of p: b = B.nbr[k * N + p], k < cnt[p], ascending
#define FOR_NBR(p, b)                                      \
  for (u32 k_ = 0, n_ = B.cnt[p]; k_ < n_; k_++)          \
    if (const u32 b = B.nbr[(size_t)k_ * P.N + (p)]; true)

template <bool kGrid>
__device__ void buildNeighbours(const Prm& P, const Buf& B, const Team<kGrid>& T) {
  const u32 N = P.N, C = P.C;
  __shared__ u32 s_sum[kThreads];
  for (u32 c = T.rank; c < C; c += T.size) B.cellCount[c] = 0u;
  T.sync();
  for (u32 p = T.rank; p < N; p += T.size) {
    const d2 x = B.x[p];
    int cx = (int)floor(x.x / P.support), cy = (int)floor(x.y / P.support);
    cx = min(max(cx, 0), P.ncx - 1);
    cy = min(max(cy, 0), P.ncy - 1);
    const u32 c = (u32)(cy * P.ncx + cx);
    B.cellOf[p] = c;
    atomicAdd(&B.cellCount[c], 1u);
  }
  T.sync();
  // exclusive scan by workgroup 0: a contiguous chunk of cells per thread, then
  // the chunk sums
  if (blockIdx.x == 0) {
    const u32 tid = threadIdx.x, nt = blockDim.x;
    const u32 chunk = (C + nt - 1) / nt;
    const u32 c0 = min(tid * chunk, C), c1 = min(c0 + chunk, C);
    u32 s = 0;
    for (u32 c = c0; c < c1; c++) s += B.cellCount[c];
    s_sum[tid] = s;
    __syncthreads();
    for (u32 off = 1; off < nt; off <<= 1) {
      const u32 add = tid >= off ? s_sum[tid - off] : 0u;
      __syncthreads();
      s_sum[tid] += add;
      __syncthreads();
    }
    u32 run = s_sum[tid] - s;
    for (u32 c = c0; c < c1; c++) {
      B.cellStart[c] = run;
      B.cursor[c] = run;
      run += B.cellCount[c];
    }
    if (tid == 0) B.cellStart[C] = N;
  }
  T.sync();
  for (u32 p = T.rank; p < N; p += T.size) B.sorted[atomicAdd(&B.cursor[B.cellOf[p]], 1u)] = p;
  T.sync();
  for (u32 p = T.rank; p < N; p += T.size) {
    const d2 xp = B.x[p];
    const int cx = (int)(B.cellOf[p] % (u32)P.ncx), cy = (int)(B.cellOf[p] / (u32)P.ncx);
    const int x0 = max(cx - 1, 0), x1 = min(cx + 1, P.ncx - 1);
    u32 n = 0;
    for (int y = max(cy - 1, 0); y <= min(cy + 1, P.ncy - 1); y++) {
      const u32 row = (u32)(y * P.ncx);
      for (u32 s2 = B.cellStart[row + x0], e = B.cellStart[row + x1 + 1]; s2 < e; s2++) {
        const u32 b = B.sorted[s2];
        if (b == p) continue;
        const d2 xb = B.x[b];
        const double d0 = xp.x - xb.x, d1 = xp.y - xb.y;
        if (!(d0 * d0 + d1 * d1 <= P.supp2)) continue;
        if (n == kMaxNbr) {
          atomicOr(B.err, 1u);
          break;
        }
        // insertion into the ascending list
        u32 k = n;
        while (k > 0 && B.nbr[(size_t)(k - 1) * N + p] > b) {
          B.nbr[(size_t)k * N + p] = B.nbr[(size_t)(k - 1) * N + p];
          k--;
        }
        B.nbr[(size_t)k * N + p] = b;
        n++;
      }
    }
    B.cnt[p] = n;
    const m2 F = B.F[p];
    B.Vn[p] = P.V0 * (F.a * F.d - F.b * F.c);
  }
  T.sync();
}

// sum_b (V_b (v_b - v_p)_i) gt_j with gt = L_p gradW(x_p - x_b) (xpbi2d.py:157-160)
__device__ __forceinline__ m2 gradV(const Prm& P, const Buf& B, u32 p, const m2& L, d2 xp, d2 vp) {
  m2 g{0.0, 0.0, 0.0, 0.0};
  FOR_NBR(p, b) {
    const d2 dxv = sub(xp, B.x[b]);
    const d2 gt = mv(L, kGradW(P, dxv, nrm(dxv)));
    const d2 dv = sub(B.v[b], vp);
    const double Vb = B.Vn[b];
    g.a = g.a + (Vb * dv.x) * gt.x;
    g.b = g.b + (Vb * dv.x) * gt.y;
    g.c = g.c + (Vb * dv.y) * gt.x;
    g.d = g.d + (Vb * dv.y) * gt.y;
  }
  return g;
}
__device__ __forceinline__ m2 trialF(const Prm& P, const m2& gv, const m2& F) {
  const m2 M{1.0 + P.dt * gv.a, 0.0 + P.dt * gv.b, 0.0 + P.dt * gv.c, 1.0 + P.dt * gv.d};
  return mm(M, F);
}

template <bool kGrid>
__global__ __launch_bounds__(kThreads) void k_run(Prm P, Buf B, int substeps) {
  const Team<kGrid> T;
  const u32 N = P.N;
  for (int step = 0; step < substeps; step++) {
    buildNeighbours(P, B, T);
    // Bonet-Lok L (xpbi2d.py:142-148), gravity, lambda = 0
    for (u32 p = T.rank; p < N; p += T.size) {
      const d2 xp = B.x[p];
      m2 A{0.0, 0.0, 0.0, 0.0};
      FOR_NBR(p, b) {
        const d2 dxv = sub(xp, B.x[b]);
        const d2 gW = kGradW(P, dxv, nrm(dxv));
        const double Vb = B.Vn[b];
        A.a = A.a + (Vb * gW.x) * (-dxv.x);
        A.b = A.b + (Vb * gW.x) * (-dxv.y);
        A.c = A.c + (Vb * gW.y) * (-dxv.x);
        A.d = A.d + (Vb * gW.y) * (-dxv.y);
      }
      const Svd s = svd2(A);
      const double thr = 1e-6 * fmax(s.s0, 1e-12);
      const double i0 = s.s0 > thr ? 1.0 / s.s0 : 0.0;
      const double i1 = s.s1 > thr ? 1.0 / s.s1 : 0.0;
      B.L[p] = usv(s.V, i0, i1, s.U);
      const d2 v = B.v[p];
      B.v[p] = d2{v.x + P.dt * P.g0, v.y + P.dt * P.g1};
      B.lam[p] = 0.0;
    }
    T.sync();
    for (int it = 0; it < P.iterations; it++) {
      // constraint of each particle (xpbi2d.py:155-182)
      for (u32 p = T.rank; p < N; p += T.size) {
        const d2 xp = B.x[p], vp = B.v[p];
        const m2 L = B.L[p], F = B.F[p];
        const Svd s = svd2(trialF(P, gradV(P, B, p, L, xp, vp), F));
        double sp0, sp1, dp0, dp1;
        druckerPrager(P, s.s0, s.s1, &sp0, &sp1);
        const double psi = stvk(P, sp0, sp1, &dp0, &dp1);
        const double C = sqrt(fmax(2.0 * psi, 0.0));
        const double Cc = fmax(C, 1e-8);
        const m2 Pk0 = usv(s.U, dp0, dp1, s.V);
        const m2 Pk{Pk0.a / Cc, Pk0.b / Cc, Pk0.c / Cc, Pk0.d / Cc};
        const m2 G{Pk.a * F.a + Pk.b * F.b, Pk.a * F.c + Pk.b * F.d, Pk.c * F.a + Pk.d * F.b,
                   Pk.c * F.c + Pk.d * F.d};  // Pk F^T
        d2 gcp{0.0, 0.0};
        double den = 0.0;
        FOR_NBR(p, b) {
          const d2 dxv = sub(xp, B.x[b]);
          const d2 gt = mv(L, kGradW(P, dxv, nrm(dxv)));
          const double Vb = B.Vn[b];
          const d2 w = mv(G, gt);
          const d2 gc{Vb * w.x, Vb * w.y};
          gcp.x = gcp.x + -gc.x;
          gcp.y = gcp.y + -gc.y;
          den = den + (gc.x * gc.x + gc.y * gc.y) / P.mass;
        }
        den = den + (gcp.x * gcp.x + gcp.y * gcp.y) / P.mass;
        const double lam = B.lam[p];
        double dl = (-C - P.alphaT * lam) / (den + P.alphaT);
        dl = C > 1e-9 ? P.omega * dl : 0.0;
        B.lam[p] = lam + dl;
        B.dlam[p] = dl;
        B.G[p] = G;
        B.gCp[p] = gcp;
      }
      T.sync();
      // Jacobi velocity update (xpbi2d.py:183-187): gathered over the pairs (b, i)
      for (u32 i = T.rank; i < N; i += T.size) {
        const d2 xi = B.x[i];
        const double Vi = B.Vn[i];
        d2 acc{0.0, 0.0};
        FOR_NBR(i, b) {
          const d2 dxv = sub(B.x[b], xi);  // x_P - x_B of the pair (b, i)
          const d2 gt = mv(B.L[b], kGradW(P, dxv, nrm(dxv)));
          const d2 w = mv(B.G[b], gt);
          const d2 gc{Vi * w.x, Vi * w.y};
          const double dl = B.dlam[b];
          acc.x = acc.x + (gc.x * dl) / P.mdt;
          acc.y = acc.y + (gc.y * dl) / P.mdt;
        }
        const d2 g = B.gCp[i];
        const double dl = B.dlam[i];
        acc.x = acc.x + (g.x * dl) / P.mdt;
        acc.y = acc.y + (g.y * dl) / P.mdt;
        const d2 v = B.v[i];
        const d2 vn{v.x + acc.x, v.y + acc.y};
        B.v[i] = vn;
        B.xc[i] = d2{xi.x + P.dt * vn.x, xi.y + P.dt * vn.y};
      }
      T.sync();
      // position correction (xpbi2d.py:189-201) and walls (xpbi2d.py:203-220)
      for (u32 i = T.rank; i < N; i += T.size) {
        const d2 xci = B.xc[i];
        d2 acc{0.0, 0.0};
        FOR_NBR(i, b) {  // pairs (i, b): +corr
          const d2 dxc = sub(xci, B.xc[b]);
          const double dc = nrm(dxc);
          const double Cd = dc - P.hc;
          if (Cd < 0.0) {
            const double m = fmax(dc, 1e-12);
            const d2 n = dc > 1e-12 ? d2{dxc.x / m, dxc.y / m} : d2{0.0, 0.0};
            acc.x = acc.x + (-0.5 * Cd) * n.x;
            acc.y = acc.y + (-0.5 * Cd) * n.y;
          } else {
            acc.x = acc.x + 0.0;
            acc.y = acc.y + 0.0;
          }
        }
        FOR_NBR(i, b) {  // pairs (b, i): -corr
          const d2 dxc = sub(B.xc[b], xci);
          const double dc = nrm(dxc);
          const double Cd = dc - P.hc;
          if (Cd < 0.0) {
            const double m = fmax(dc, 1e-12);
            const d2 n = dc > 1e-12 ? d2{dxc.x / m, dxc.y / m} : d2{0.0, 0.0};
            acc.x = acc.x + -((-0.5 * Cd) * n.x);
            acc.y = acc.y + -((-0.5 * Cd) * n.y);
          } else {
            acc.x = acc.x + -0.0;
            acc.y = acc.y + -0.0;
          }
        }
        const d2 xi = B.x[i];
        d2 v = B.v[i];
        v.x = v.x + (acc.x / P.dt) * 0.5;
        v.y = v.y + (acc.y / P.dt) * 0.5;
        const d2 xw{xi.x + P.dt * v.x, xi.y + P.dt * v.y};
        double vv[2] = {v.x, v.y};
        const double xx[2] = {xi.x, xi.y}, xwv[2] = {xw.x, xw.y}, dom[2] = {P.dom0, P.dom1};
#pragma unroll
        for (int d = 0; d < 2; d++) {
          const int td = 1 - d;
          const bool lo = xwv[d] < P.h, hi = xwv[d] > dom[d] - P.h;
#pragma unroll
          for (int side = 0; side < 2; side++) {
            if (!(side == 0 ? lo : hi)) continue;
            const double target = side == 0 ? P.h : dom[d] - P.h;
            const double vn = (target - xx[d]) / P.dt;
            const double dvn = vn - vv[d];
            const double vt = vv[td];
            const double fric = fmax(0.0, 1.0 - (0.4 * fabs(dvn)) / (fabs(vt) + 1e-9));
            vv[td] = vt * fric;
            vv[d] = vn;
          }
        }
        B.v[i] = d2{vv[0], vv[1]};
      }
      T.sync();
    }
    // XSPH (xpbi2d.py:222-225)
    for (u32 p = T.rank; p < N; p += T.size) {
      const d2 xp = B.x[p], vp = B.v[p];
      d2 acc{0.0, 0.0};
      FOR_NBR(p, b) {
        const d2 dxv = sub(xp, B.x[b]);
        const double W = kW(P, nrm(dxv));
        const d2 dv = sub(B.v[b], vp);
        const double Vb = B.Vn[b];
        acc.x = acc.x + (Vb * dv.x) * W;
        acc.y = acc.y + (Vb * dv.y) * W;
      }
      B.dvx[p] = acc;
    }
    T.sync();
    for (u32 p = T.rank; p < N; p += T.size) {
      const d2 v = B.v[p], a = B.dvx[p];
      B.v[p] = d2{v.x + P.xsphC * a.x, v.y + P.xsphC * a.y};
    }
    T.sync();
    // F update with the return map (xpbi2d.py:227-233)
    for (u32 p = T.rank; p < N; p += T.size) {
      const m2 F = B.F[p];
      const Svd s = svd2(trialF(P, gradV(P, B, p, B.L[p], B.x[p], B.v[p]), F));
      double sp0, sp1;
      druckerPrager(P, s.s0, s.s1, &sp0, &sp1);
      B.F[p] = usv(s.U, sp0, sp1, s.V);
    }
    T.sync();
    // positions (xpbi2d.py:235-237)
    for (u32 p = T.rank; p < N; p += T.size) {
      const d2 x = B.x[p], v = B.v[p];
      const double lo = 0.5 * P.h;
      B.x[p] = d2{fmin(fmax(x.x + P.dt * v.x, lo), P.dom0 - 0.5 * P.h),
                  fmin(fmax(x.y + P.dt * v.y, lo), P.dom1 - 0.5 * P.h)};
    }
    T.sync();
  }
}

}  // namespace xpbi2d

// ---------------------------------------------------------------- C ABI

using namespace xpbi2d;

struct xpbi2d_sim {
  Prm P{};
  Buf B{};
  int dev = 0;
  int single = 0;     // 1: one workgroup; 0: a cooperative grid, one thread per particle
  int coopBlocks = 0;  // blocks of the cooperative grid (capped by co-residency)
  hipStream_t stream = nullptr;
  std::string err;
  void* mem = nullptr;
};

static int fail2(xpbi2d_sim* s, int code, const std::string& m) {
  if (s) s->err = m;
  return code;
}
static int hipfail2(xpbi2d_sim* s, hipError_t e, const char* what) {
  return fail2(s, XPBI_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

extern "C" {

int xpbi2d_create(uint32_t N, const double* positions, const xpbi2d_params* p, int device, xpbi2d_sim** out) {
  if (!out) return XPBI_ERR_ARG;
  *out = nullptr;
  if (!p || (N > 0 && !positions)) return XPBI_ERR_ARG;
  if (N == 0 || N > XPBI2D_MAX_PARTICLES) return XPBI_ERR_ARG;
  if (!(p->h > 0.0) || !(p->dt > 0.0) || p->iterations < 0 || !(p->domain[0] > 2.0 * p->h) ||
      !(p->domain[1] > 2.0 * p->h) || !(p->E > 0.0) || !(p->rho > 0.0) || !(p->nu > -1.0 && p->nu < 0.5))
    return XPBI_ERR_ARG;
  auto* s = new xpbi2d_sim();
  s->dev = device;
  Prm& P = s->P;
  const double h = p->h;
  P.h = h;
  P.support = 2.0 * h;
  P.supp2 = P.support * P.support;
  P.dt = p->dt;
  P.V0 = h * h;
  P.mass = p->rho * P.V0;
  P.mu = p->E / (2 * (1 + p->nu));
  P.lam = p->E * p->nu / ((1 + p->nu) * (1 - 2 * p->nu));
  const double phi = p->phi_deg * (M_PI / 180.0);
  P.alphaDP = std::sqrt(2.0 / 3.0) * 2.0 * std::sin(phi) / (3.0 - std::sin(phi));
  P.wsig = 7.0 / (4.0 * M_PI * h * h);
  P.xsphC = p->xsph_c;
  P.omega = p->omega;
  P.alphaT = (1.0 / P.V0) / (P.dt * P.dt);
  P.mdt = P.mass * P.dt;
  P.dom0 = p->domain[0];
  P.dom1 = p->domain[1];
  P.g0 = p->gravity[0];
  P.g1 = p->gravity[1];
  P.hc = h - 0.25 * h;
  P.iterations = p->iterations;
  P.ncx = (int)std::ceil(P.dom0 / P.support);
  P.ncy = (int)std::ceil(P.dom1 / P.support);
  P.N = N;
  P.C = (u32)P.ncx * (u32)P.ncy;
  if ((double)P.ncx * P.ncy > 16.0e6) {
    delete s;
    return XPBI_ERR_ARG;
  }
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete s;
    return XPBI_ERR_HIP;
  }
  // one allocation, 256-B aligned slices
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t oX = take(16ull * N), oV = take(16ull * N), oXc = take(16ull * N), oGcp = take(16ull * N),
               oDvx = take(16ull * N), oF = take(32ull * N), oL = take(32ull * N), oG = take(32ull * N),
               oVn = take(8ull * N), oLam = take(8ull * N), oDl = take(8ull * N), oCc = take(4ull * (P.C + 1)),
               oCs = take(4ull * (P.C + 1)), oCu = take(4ull * (P.C + 1)), oSo = take(4ull * N),
               oCo = take(4ull * N), oNb = take(4ull * kMaxNbr * N), oCn = take(4ull * N), oEr = take(4);
  if ((e = hipMalloc(&s->mem, off)) != hipSuccess) {
    delete s;
    return XPBI_ERR_NOMEM;
  }
  char* m = (char*)s->mem;
  Buf& B = s->B;
  B.x = (d2*)(m + oX);
  B.v = (d2*)(m + oV);
  B.xc = (d2*)(m + oXc);
  B.gCp = (d2*)(m + oGcp);
  B.dvx = (d2*)(m + oDvx);
  B.F = (m2*)(m + oF);
  B.L = (m2*)(m + oL);
  B.G = (m2*)(m + oG);
  B.Vn = (double*)(m + oVn);
  B.lam = (double*)(m + oLam);
  B.dlam = (double*)(m + oDl);
  B.cellCount = (u32*)(m + oCc);
  B.cellStart = (u32*)(m + oCs);
  B.cursor = (u32*)(m + oCu);
  B.sorted = (u32*)(m + oSo);
  B.cellOf = (u32*)(m + oCo);
  B.nbr = (u32*)(m + oNb);
  B.cnt = (u32*)(m + oCn);
  B.err = (u32*)(m + oEr);
  if ((e = hipStreamCreate(&s->stream)) != hipSuccess) {
    hipFree(s->mem);
    delete s;
    return XPBI_ERR_HIP;
  }
  // x from the caller; v = 0, F = I, lambda = 0 (xpbi2d.py:52-58)
  std::string hostInit(off, '\0');
  char* hm = &hostInit[0];
  std::memcpy(hm + oX, positions, 16ull * N);
  m2* Fh = (m2*)(hm + oF);
  for (u32 i = 0; i < N; i++) Fh[i] = m2{1.0, 0.0, 0.0, 1.0};
  if ((e = hipMemcpy(s->mem, hm, off, hipMemcpyHostToDevice)) != hipSuccess) {
    hipStreamDestroy(s->stream);
    hipFree(s->mem);
    delete s;
    return XPBI_ERR_HIP;
  }
  // the cooperative grid: one 256-thread workgroup per 256 particles, no more
  // than can be resident at once (the grid barrier needs every block running)
  int coop = 0, cus = 0, perCU = 0;
  hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, (const void*)k_run<true>, 256, 0);
  const int want = (int)((N + 255) / 256);
  s->coopBlocks = want < perCU * cus ? want : perCU * cus;
  s->single = (coop && s->coopBlocks > 0) ? 0 : 1;
  *out = s;
  return XPBI_OK;
}

int xpbi2d_set_workgroups(xpbi2d_sim* s, int single) {
  if (!s) return XPBI_ERR_ARG;
  if (single != 0 && single != 1) return fail2(s, XPBI_ERR_ARG, "single must be 0 or 1");
  if (!single && s->coopBlocks <= 0) return fail2(s, XPBI_ERR_STATE, "cooperative launch unavailable");
  s->single = single;
  return XPBI_OK;
}

void xpbi2d_destroy(xpbi2d_sim* s) {
  if (!s) return;
  hipSetDevice(s->dev);
  if (s->stream) hipStreamSynchronize(s->stream);
  if (s->mem) hipFree(s->mem);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
}

const char* xpbi2d_last_error(xpbi2d_sim* s) { return s ? s->err.c_str() : "null handle"; }

static int run2d(xpbi2d_sim* s, int substeps, float* ms) {
  if (!s) return XPBI_ERR_ARG;
  if (substeps < 0) return fail2(s, XPBI_ERR_ARG, "substeps < 0");
  hipSetDevice(s->dev);
  hipEvent_t a = nullptr, b = nullptr;
  if (ms) {
    hipEventCreate(&a);
    hipEventCreate(&b);
    hipEventRecord(a, s->stream);
  }
  hipError_t e = hipSuccess;
  if (substeps > 0) {
    if (s->single) {
      hipLaunchKernelGGL(k_run<false>, dim3(1), dim3(kThreads), 0, s->stream, s->P, s->B, substeps);
    } else {
      void* args[] = {(void*)&s->P, (void*)&s->B, (void*)&substeps};
      e = hipLaunchCooperativeKernel((const void*)k_run<true>, dim3(s->coopBlocks), dim3(256), args, 0, s->stream);
      if (e != hipSuccess) return hipfail2(s, e, "cooperative launch of k_run");
    }
  }
  e = hipGetLastError();
  if (e != hipSuccess) return hipfail2(s, e, "k_run launch");
  if (ms) hipEventRecord(b, s->stream);
  if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return hipfail2(s, e, "hipStreamSynchronize");
  if (ms) {
    hipEventElapsedTime(ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
  }
  u32 err = 0;
  if ((e = hipMemcpy(&err, s->B.err, 4, hipMemcpyDeviceToHost)) != hipSuccess) return hipfail2(s, e, "hipMemcpy");
  if (err) return fail2(s, XPBI_ERR_STATE, "a particle has more than 48 neighbours within 2h; state is invalid");
  return XPBI_OK;
}

int xpbi2d_step(xpbi2d_sim* s, int substeps) { return run2d(s, substeps, nullptr); }

int xpbi2d_time(xpbi2d_sim* s, int substeps, float* ms) {
  if (!ms) return fail2(s, XPBI_ERR_ARG, "ms is null");
  return run2d(s, substeps, ms);
}

static void* field2d(xpbi2d_sim* s, int field, size_t* bytes) {
  const size_t N = s->P.N;
  switch (field) {
    case XPBI2D_FIELD_X: *bytes = 16 * N; return s->B.x;
    case XPBI2D_FIELD_V: *bytes = 16 * N; return s->B.v;
    case XPBI2D_FIELD_F: *bytes = 32 * N; return s->B.F;
    case XPBI2D_FIELD_LAMBDA: *bytes = 8 * N; return s->B.lam;
    default: *bytes = 0; return nullptr;
  }
}

int xpbi2d_read(xpbi2d_sim* s, int field, double* dst, size_t count) {
  if (!s) return XPBI_ERR_ARG;
  size_t bytes;
  void* src = field2d(s, field, &bytes);
  if (!src) return fail2(s, XPBI_ERR_ARG, "unknown field");
  if (count * sizeof(double) != bytes) return fail2(s, XPBI_ERR_ARG, "count does not match the field size");
  if (!dst) return fail2(s, XPBI_ERR_ARG, "dst is null");
  hipSetDevice(s->dev);
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? XPBI_OK : hipfail2(s, e, "read");
}

int xpbi2d_write(xpbi2d_sim* s, int field, const double* src, size_t count) {
  if (!s) return XPBI_ERR_ARG;
  if (field == XPBI2D_FIELD_LAMBDA) return fail2(s, XPBI_ERR_ARG, "lambda is reset every substep; read only");
  size_t bytes;
  void* dst = field2d(s, field, &bytes);
  if (!dst) return fail2(s, XPBI_ERR_ARG, "unknown field");
  if (count * sizeof(double) != bytes) return fail2(s, XPBI_ERR_ARG, "count does not match the field size");
  if (!src) return fail2(s, XPBI_ERR_ARG, "src is null");
  hipSetDevice(s->dev);
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? XPBI_OK : hipfail2(s, e, "write");
}

}  // extern "C"
