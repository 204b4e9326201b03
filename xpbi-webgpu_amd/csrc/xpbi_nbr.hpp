// Neighbour iteration (cell rows, support-neighbour list, batched gathers) and
// the per-particle finishing code of the constraint sweeps.
#pragma once
#include "xpbi_internal.hpp"

namespace xpbi {

constexpr u32 kRowCap = 384u;  // simkernels.js:73

// 9 row spans [cellStart[c(x-1)], cellStart[c(x+1)+1]) capped at beg+384;
// cellStart[C] = N replaces the reference's end-of-table branch.
template <class Fn>
__device__ __forceinline__ void forRows(const KP& P, const u32* __restrict__ cs, i3 cc, Fn&& fn) {
  const u32 x0 = (u32)max(cc.x - 1, 0), x1 = (u32)min(cc.x + 1, P.gx - 1);
#pragma unroll 1
  for (int dz = -1; dz <= 1; dz++) {
    int z = cc.z + dz;
    if (z < 0 || z >= P.gz) continue;
#pragma unroll 1
    for (int dy = -1; dy <= 1; dy++) {
      int y = cc.y + dy;
      if (y < 0 || y >= P.gy) continue;
      u32 row = P.cellBase + (u32)y * (u32)P.gx + (u32)z * P.gxy;
      u32 beg = cs[row + x0];
      u32 end = min(cs[row + x1 + 1u], beg + kRowCap);
#pragma unroll 1
      for (u32 b = beg; b < end; b++) fn(b);
    }
  }
}

// the same 9 spans as forRows, handed over whole: fn(beg, end)
template <class Fn>
__device__ __forceinline__ void forRowSpans(const KP& P, const u32* __restrict__ cs, i3 cc, Fn&& fn) {
  const u32 x0 = (u32)max(cc.x - 1, 0), x1 = (u32)min(cc.x + 1, P.gx - 1);
#pragma unroll 1
  for (int dz = -1; dz <= 1; dz++) {
    int z = cc.z + dz;
    if (z < 0 || z >= P.gz) continue;
#pragma unroll 1
    for (int dy = -1; dy <= 1; dy++) {
      int y = cc.y + dy;
      if (y < 0 || y >= P.gy) continue;
      u32 row = P.cellBase + (u32)y * (u32)P.gx + (u32)z * P.gxy;
      u32 beg = cs[row + x0];
      fn(beg, min(cs[row + x1 + 1u], beg + kRowCap));
    }
  }
}

// forRows with the candidates' records loaded kB at a time (all loads of a batch
// in flight together); use(b, record) in candidate order
template <int kB, class Load, class Use>
__device__ __forceinline__ void forRowsG(const KP& P, const u32* __restrict__ cs, i3 cc, Load&& load, Use&& use) {
  forRowSpans(P, cs, cc, [&](u32 beg, u32 end) {
    for (u32 b0 = beg; b0 < end; b0 += (u32)kB) {
      decltype(load(0u)) d[kB];
#pragma unroll
      for (int j = 0; j < kB; j++)
        if (b0 + j < end) d[j] = load(b0 + j);
#pragma unroll
      for (int j = 0; j < kB; j++)
        if (b0 + j < end) use(b0 + j, d[j]);
    }
  });
}

// Support-neighbour list.  Entry k of particle i is lst[k * N + i], the u16
// offset b - head[i].x of neighbour b from the particle's first candidate
// (cellStart at the stencil's lowest cell (x-1, y-1, z-1), clamped to the grid):
// the candidates of a particle span two z-layers of cells and a little more.
// head[i].y is the neighbour count; head[i].y > K (more than K neighbours, or
// kNbrFlag set for an offset past 16 bits) makes the particle's sweeps walk the
// cell rows.
typedef uint16_t u16;
constexpr u32 kNbrFlag = 0x80000000u;
constexpr u32 kNbrOffMax = 0xffffu;
struct NLc {
  const uint2* __restrict__ head;
  const u16* __restrict__ lst;
  u32 K, N;
};

#ifndef XPBI_GATHER_BATCH
#define XPBI_GATHER_BATCH 8
#endif
// the constraint gather of solveB carries 17 floats per neighbour: its own depth
#ifndef XPBI_GATHER_BATCH_B
#define XPBI_GATHER_BATCH_B 6
#endif
constexpr int kBatch = XPBI_GATHER_BATCH;
constexpr int kBatchB = XPBI_GATHER_BATCH_B;

// Support-neighbour walk with the gathers issued in batches of kBatch
// independent loads (and the next batch's indices prefetched) so that several
// cache misses are in flight per lane.  load(b) gathers the neighbour's record,
// whose member p is (x_b, tag_b); use(b, rec, dxv = x_i - x_b, |dxv|^2)
// consumes it, in list order.
template <int kB = kBatch, class Load, class Use>
__device__ __forceinline__ void forSupportG(const KP& P, const u32* __restrict__ cs, const NLc& nl, u32 i, f3 xi,
                                            Load&& load, Use&& use) {
  const uint2 hd = nl.head[i];
  const u32 n = hd.y;
  if (n <= nl.K) {
    // raw entries of the next batch (past the count: the particle's own offset,
    // so a batch's tail gathers i itself); the base is added when the batch is
    // used, so that no arithmetic on a loaded entry sits inside the load's
    // predicate (which would wait on every load in flight).  Slot kk's row
    // nl.lst + kk N is uniform: the entry load is a scalar base plus 2 i.  (The
    // predicate stays: unpredicated entry loads, over a list padded with the own
    // offset, made the compiler wait for a whole batch of gathers before its
    // first use and measured 13 % slower.)
    const u32 self = i - hd.x;
    u32 en[kB];
#pragma unroll
    for (int j = 0; j < kB; j++) en[j] = ((u32)j < n) ? (u32)gat(nl.lst + (size_t)j * nl.N, i) : self;
    for (u32 k = 0; k < n; k += kB) {
      u32 b[kB];
      decltype(load(0u)) d[kB];
#pragma unroll
      for (int j = 0; j < kB; j++) {
        b[j] = hd.x + en[j];
        d[j] = load(b[j]);
      }
#pragma unroll
      for (int j = 0; j < kB; j++) {
        u32 kk = k + kB + (u32)j;
        en[j] = (kk < n) ? (u32)gat(nl.lst + (size_t)kk * nl.N, i) : self;
      }
#pragma unroll
      for (int j = 0; j < kB; j++) {
        if (k + (u32)j < n) {
          const f3 dd = xi - xyz(d[j].p);
          use(b[j], d[j], dd, dot(dd, dd));
        }
      }
    }
  } else {
    forRows(P, cs, cellCoord(P, xi), [&](u32 b) {
      if (b != i) {
        const auto rec = load(b);
        const f3 d = xi - xyz(rec.p);
        const float r2 = dot(d, d);
        if (r2 < P.supp2) use(b, rec, d, r2);
      }
    });
  }
}

// neighbour record of the velocity sweeps: position/tag and velocity/volume
struct PV {
  float4 p, v;
};

// material tables staged in LDS at kernel entry
struct MatLds {
  DevMat m[8];
  float mass[8];
  float invMass[8];
  float wc[8][8];  // contact weight m_b / (m_a + m_b) of material pair (a, b) (simkernels.js:346)
  float mRho[8][8];  // mass of material b over density of material a (simkernels.js:429, :470)
  u32 solidMask;   // bit k set when material k is not water
};
// every material of the scene is solid: the per-neighbour phase tests vanish
__device__ __forceinline__ bool allSolid(const MatLds& t) { return t.solidMask == 0xFFu; }
__device__ __forceinline__ void loadMatLds(MatLds& t, const DevMat* __restrict__ mats) {
  const u32 tid = threadIdx.x;
  __syncthreads();  // a previous scene's table may still be read
  if (tid < 128) ((float*)t.m)[tid] = ((const float*)mats)[tid];
  if (tid < 8) {
    t.mass[tid] = mats[tid].mass;
    t.invMass[tid] = 1.0f / mats[tid].mass;
  }
  if (tid < 64) {
    const float ma = mats[tid >> 3].mass, mb = mats[tid & 7u].mass;
    t.wc[tid >> 3][tid & 7u] = mb / (ma + mb);
    t.mRho[tid >> 3][tid & 7u] = mb / mats[tid >> 3].density;
  }
  if (tid == 0) {
    u32 mask = 0;
    for (int k = 0; k < 8; k++)
      if (mats[k].kind != MAT_WATER) mask |= 1u << k;
    t.solidMask = mask;
  }
  __syncthreads();
}
// the scene's 8 materials (mats holds 8 per scene), reloaded when the scene changes
__device__ __forceinline__ void sceneMats(MatLds& t, const DevMat* __restrict__ mats, u32 scene, u32* cur) {
  if (scene != *cur) {
    loadMatLds(t, mats + 8 * scene);
    *cur = scene;
  }
}
__device__ __forceinline__ MatV matv(const MatLds& t, u32 k) {
  const DevMat& r = t.m[k];
  return MatV{r.mu, r.lam, r.density, r.kind, r.p0, r.p1, r.p2, r.p3, r.boundFriction, r.mass};
}
__device__ __forceinline__ bool solidTag(u32 mask, float w) { return (mask >> matOfTag(w)) & 1u; }

// ---------------------------------------------------------------- velocity gradient
// simkernels.js:224-245.  With g_b = gradW(x_i - x_b) and w_b = V_b (v_b - v_i),
//   grad v = sum_b w_b (x) (L_i g_b) = (sum_b w_b (x) g_b) L_i^T = M L_i^T,
// so the sweep accumulates the first moment M and L_i is applied once.
struct Moments {
  f33 M;    // sum_b V_b (v_b - v_i) (x) g_b
  f3 s;     // sum_b V_b g_b
  float q00, q11, q22, q01, q02, q12;  // sum_b (V_b^2 / m_b) g_b g_b^T
};
__device__ __forceinline__ void momentsZero(Moments& m) {
  m.M = mzero<float>();
  m.s = splat3(0.0f);
  m.q00 = m.q11 = m.q22 = m.q01 = m.q02 = m.q12 = 0.0f;
}
// one neighbour's contribution: g = gradW(x_i - x_b), w = V_b (v_b - v_i)
template <bool kWithQ>
__device__ __forceinline__ void momentsAdd(Moments& m, f3 g, f3 w, float Vb, float invMb) {
  m.M = madd(m.M, outer(w, g));
  if (kWithQ) {
    m.s = m.s + g * Vb;
    const float c = Vb * Vb * invMb;
    const f3 cg = g * c;
    m.q00 += cg.x * g.x;
    m.q11 += cg.y * g.y;
    m.q22 += cg.z * g.z;
    m.q01 += cg.x * g.y;
    m.q02 += cg.x * g.z;
    m.q12 += cg.y * g.z;
  }
}
// kAllSolid: the scene has no water, every neighbour contributes
template <bool kWithQ, bool kAllSolid>
__device__ __forceinline__ Moments neighbourMoments(const KP& P, const u32* __restrict__ cs,
                                                    const float4* __restrict__ pos, const float4* __restrict__ vin,
                                                    const NLc& nl, const float* invMass, u32 solidMask, u32 i,
                                                    f3 xi, f3 vi) {
  Moments m;
  momentsZero(m);
  forSupportG(P, cs, nl, i, xi, [&](u32 b) { return PV{gat(pos, b), gat(vin, b)}; },
              [&](u32 b, const PV& r, f3 dxv, float r2) {
                const float4 pb = r.p, vb = r.v;
                if (kAllSolid || solidTag(solidMask, pb.w)) {
                  momentsAdd<kWithQ>(m, gradWfast(P, dxv, r2), (xyz(vb) - vi) * vb.w, vb.w,
                                     invMass[matOfTag(pb.w)]);
                }
              });
  return m;
}

// Frobenius product A : B of two column-major matrices
__device__ __forceinline__ float frob(const f33& A, const f33& B) {
  return dot(A.c[0], B.c[0]) + dot(A.c[1], B.c[1]) + dot(A.c[2], B.c[2]);
}

__device__ __forceinline__ void blockReduce2(double a, double b, double* __restrict__ out) {
  __shared__ double s_a[4], s_b[4];
  const u32 lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  if (lane == 0) { s_a[w] = a; s_b[w] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = (s_a[0] + s_a[1]) + (s_a[2] + s_a[3]);
    out[1] = (s_b[0] + s_b[1]) + (s_b[2] + s_b[3]);
  }
}


// simkernels.js:259-300 after the neighbour sweep, for a solid particle: trial
// F, SVD, return map, C = sqrt(2 Psi) and, when C >= 1e-7, GL, sgrad, the
// Delta-lambda denominator and dC.v from the moments.
struct Constraint {
  float C, denom, cdot;
  f33 GL;
  f3 sgrad;
  f3 sigTrial, sigPost;
  float hbT;  // HB root position of the return map (NaN when no bisection ran)
  bool active;
};
__device__ __forceinline__ Constraint evalConstraint(const KP& P, const MatLds& T, u32 i, float4 pi, float4 a,
                                                     const Moments& mo, M3C Fm, M3C Lm, float hbHint,
                                                     bool wantHbT) {
  Constraint c;
  const u32 mi = matOfTag(pi.w);
  MatV m = hardenedMat(matv(T, mi), a.z);
  const f33 Fn = ldm(Fm, i);
  const f33 Li = ldm(Lm, i);
  const f33 gradV = mul(mo.M, transpose(Li));
  f33 Ftr = mul(madd(mident<float>(), mscale(gradV, P.dt)), Fn);
  SVD3<float> sv = svd3_guarded(Ftr);
  float st;
  c.hbT = __builtin_nanf("");
  f3 sig = returnMap(sv.sig, m, a.z, P.dt, &st, hbHint, wantHbT ? &c.hbT : nullptr);
  c.sigTrial = sv.sig;
  c.sigPost = sig;
  float psi = stvkPsi(sig, m.mu, m.lam);
  c.C = sqrtf(wmax(2.0f * psi, 0.0f));
  c.active = c.C >= 1e-7f;
  if (c.active) {
    f3 dpsi = stvkDPsi(sig, m.mu, m.lam);
    f33 Pk = diagMul(sv.U, dpsi / c.C, sv.V);
    c.GL = mul(mul(Pk, transpose(Fn)), Li);
    // the reference's second sweep (simkernels.js:284-299) from the moments:
    //   g'_b = V_b GL g_b, sgrad = -sum g'_b = -GL s,
    //   sum |g'_b|^2 / m_b = (GL^T GL) : Q,  cdot = sum g'_b.(v_b - v_i) = GL : M
    const f33& GL = c.GL;
    c.sgrad = -mul(GL, mo.s);
    const float h00 = dot(GL.c[0], GL.c[0]), h11 = dot(GL.c[1], GL.c[1]), h22 = dot(GL.c[2], GL.c[2]);
    const float h01 = dot(GL.c[0], GL.c[1]), h02 = dot(GL.c[0], GL.c[2]), h12 = dot(GL.c[1], GL.c[2]);
    float denom = h00 * mo.q00 + h11 * mo.q11 + h22 * mo.q22 + 2.0f * (h01 * mo.q01 + h02 * mo.q02 + h12 * mo.q12);
    c.denom = denom + dot(c.sgrad, c.sgrad) / m.mass;
    c.cdot = frob(GL, mo.M);
  }
  return c;
}

// simkernels.js:259-308: the damped XPBD update of a solid particle from the
// moments.  Writes Sp, selfDv, lambda; accumulates the residual numerator^2 and
// C^2 into *r2, *c2.
__device__ __forceinline__ void solveAFinish(const KP& P, const MatLds& T, u32 i, float4 pi, f3 vi, bool solid,
                                             const Moments& mo, M3C Fm, M3C Lm, float4* __restrict__ aux, M3P Sp,
                                             float4* __restrict__ selfDv, float* r2, float* c2,
                                             float hbHint = -1.0f) {
  bool wrote = false;
  if (solid) {
    const float4 a = aux[i];
    const Constraint c = evalConstraint(P, T, i, pi, a, mo, Fm, Lm, hbHint, false);
    if (c.active) {
      const float g = P.kDamp;
      const float num = -c.C - P.alphaT * a.y - g * P.dt * c.cdot;
      const float dlam = P.omega * num / ((1.0f + g) * c.denom + P.alphaT);
      aux[i].y = a.y + dlam;
      stm(Sp, i, mscale(c.GL, dlam));
      f3 sd = c.sgrad * dlam;
      selfDv[i] = make_float4(sd.x, sd.y, sd.z, 0.0f);
      wrote = true;
      *r2 += num * num;
      *c2 += c.C * c.C;
    }
  }
  if (!wrote) {
    stm(Sp, i, mzero<float>());
    selfDv[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

// ---------------------------------------------------------------- solveB + boundary
// simkernels.js:369-404
__device__ __forceinline__ f3 applyBoundary(const KP& P, f3 xcIn, f3 vIn, f3 xn, float fric) {
  f3 vv = vIn, xq = xcIn;
  const float lo3[3] = {P.dminx, P.dminy, P.dminz};
  const float hi3[3] = {P.dmaxx, P.dmaxy, P.dmaxz};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float lo = lo3[k], hi = hi3[k];
    float xk = comp(xq, k);
    if (xk < lo || xk > hi) {
      float bound = (xk < lo) ? lo : hi;
      float vnNew = (bound - comp(xn, k)) / P.dt;
      float dvn = fabsf(vnNew - comp(vv, k));
      f3 vt = vv;
      setc(vt, k, 0.0f);
      float vtl = len(vt);
      if (vtl > 1e-9f) {
        float drop = wmin(fric * dvn, vtl);
        vv = vv - vt / vtl * drop;
      }
      setc(vv, k, vnNew);
      xq = xn + P.dt * vv;
    }
  }
  if (P.cActive > 0.5f) {
    f3 d = xq - f3{P.cpx, P.cpy, P.cpz};
    float dist = len(d);
    if (dist < P.cR && dist > 1e-9f) {
      f3 n = d / dist;
      float vOut = dot(vv - f3{P.cvx, P.cvy, P.cvz}, n);
      float vGoal = wmin((P.cR - dist) / P.dt, 0.6f);
      if (vOut < vGoal) vv = vv + n * (vGoal - vOut);
    }
  }
  return vv;
}

struct SpGather {
  float4 p, v, a, b;
  float c;
};

// simkernels.js:338-355 for one neighbour: constraint gather dv += V_i S_b gradW(x_b - x_i)
// and the 0.75h contact on predicted positions
template <bool kAllSolid>
__device__ __forceinline__ void solveBPair(const KP& P, const MatLds& T, float Vi, const float* wc_i, f3 xci, f3 dxv,
                                           float r2, float4 pb4, float4 vb, const f33& Spb, f3* dv, f3* corr) {
  if (kAllSolid || solidTag(T.solidMask, pb4.w)) {
    const f3 dxp = -dxv;  // x_b - x_i
    *dv = *dv + Vi * mul(Spb, gradWfast(P, dxp, r2));
    const f3 xcb = xyz(pb4) + P.dt * xyz(vb);
    const f3 dxc = xci - xcb;
    const float d2 = dot(dxc, dxc);
    const float minD = P.minD;
    if (d2 < minD * minD && d2 > 1e-18f) {
      float idist = xrsq(d2);
      float dist = d2 * idist;
      float w = wc_i[matOfTag(pb4.w)];
      *corr = *corr + dxc * (idist * ((minD - dist) * w));
    }
  }
}

template <bool kAllSolid>
__device__ __forceinline__ void solveBGatherGlobal(const KP& P, const MatLds& T, u32 i, float4 pi, float4 vi4,
                                                   const float4* __restrict__ pos, const float4* __restrict__ vin,
                                                   M3C Sp, const float4* __restrict__ selfDv,
                                                   const u32* __restrict__ cs, const NLc& nl, f3* dv, f3* corr) {
  *dv = xyz(selfDv[i]);
  const float Vi = vi4.w;
  const f3 xi = xyz(pi);
  const f3 xci = xi + P.dt * xyz(vi4);
  const float* wc_i = T.wc[matOfTag(pi.w)];
  // indexed by pointer, not by gat(): the compiler then issues a batch's loads
  // grouped by array, which measured 30% faster for these five arrays than the
  // per-neighbour interleaving it emits for 32-bit offsets
  forSupportG<kBatchB>(P, cs, nl, i, xi, [&](u32 b) { return SpGather{pos[b], vin[b], Sp.a[b], Sp.b[b], Sp.c[b]}; },
              [&](u32 b, const SpGather& g, f3 dxv, float r2) {
                const f33 Spb{{f3{g.a.x, g.a.y, g.a.z}, f3{g.a.w, g.b.x, g.b.y}, f3{g.b.z, g.b.w, g.c}}};
                solveBPair<kAllSolid>(P, T, Vi, wc_i, xci, dxv, r2, g.p, g.v, Spb, dv, corr);
              });
}

// simkernels.js:357-365: impulse cap, contact relaxation, boundaries; with xcOut
// also the predicted position x + dt v the PBF sweeps of the iteration test
__device__ __forceinline__ void solveBFinish(const KP& P, const MatLds& T, u32 i, float4 pi, float4 vi4, f3 dv,
                                             f3 corr, float4* __restrict__ vout, float4* __restrict__ xcOut) {
  const u32 mi = matOfTag(pi.w);
  const float mass_i = T.mass[mi];
  const f3 xi = xyz(pi);
  f3 dvv = dv / (mass_i * P.dt);
  float dvl = len(dvv);
  if (dvl > P.dvCap) dvv = dvv * (P.dvCap / dvl);
  f3 v = xyz(vi4) + dvv + corr * (P.contactOmega / P.dt);
  v = applyBoundary(P, xi + P.dt * v, v, xi, T.m[mi].boundFriction);
  vout[i] = make_float4(v.x, v.y, v.z, vi4.w);
  if (xcOut) {
    const f3 xc = xi + P.dt * v;
    xcOut[i] = make_float4(xc.x, xc.y, xc.z, pi.w);
  }
}

}  // namespace xpbi
