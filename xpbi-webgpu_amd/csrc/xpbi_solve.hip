// Neighbour-gather kernels of the XPBI substep (gfx950).
//
// Restates the compute entry points of web/js/shaders/simkernels.js:189-542 in
// the dispatch order of web/js/sim.js:255-266:
//   computeL -> [solveA -> solveB (-> pbfA -> pbfB)] x I -> xsph -> updateF -> integrate
// One thread per particle in cell order (256-thread blocks, XCD-contiguous block
// mapping).  Positions are fixed for the whole substep, so the support
// neighbours (b != i, |x_i - x_b| < 2h, in the reference's candidate order over
// the 9 row-merged spans of simkernels.js:62-74) are found once, by computeL,
// and stored as a slot-major list; every later sweep walks that list instead of
// re-testing the ~216 candidates of the 27-cell stencil.
#include "xpbi_mlp.hpp"
#include <atomic>
#include <chrono>
#include <type_traits>

#include "xpbi_nbr.hpp"

namespace xpbi {

// lanes per particle of each sweep's neighbour walk (forSupport: 4 or 8)
#ifndef XPBI_LANES_A
#define XPBI_LANES_A 4
#endif
#ifndef XPBI_LANES_B
#define XPBI_LANES_B 4
#endif
#ifndef XPBI_LANES_X
#define XPBI_LANES_X 4
#endif
#ifndef XPBI_LANES_F
#define XPBI_LANES_F 4
#endif
constexpr u32 kLanesA = XPBI_LANES_A;  // solveA and the warm-start features pass
constexpr u32 kLanesB = XPBI_LANES_B;  // solveB
constexpr u32 kLanesX = XPBI_LANES_X;  // xsph
constexpr u32 kLanesF = XPBI_LANES_F;  // updateF

// ---------------------------------------------------------------- computeL
// simkernels.js:189-221.  A is accumulated in f32; the SVD pseudo-inverse with
// absolute tolerance 0.2 runs in FP64.  The same candidate walk records the
// support-neighbour list of the substep.
template <bool kPersist>
__global__ __launch_bounds__(256) void k_computeL(const KP* __restrict__ kps, WorkTab W, const DevMat* __restrict__ mats,
                                                  const float4* __restrict__ pos, const float4* __restrict__ vin,
                                                  const u32* __restrict__ cs, M3P L, u16* __restrict__ nLst,
                                                  uint2* __restrict__ nHead, u32 K, u32 NT,
                                                  u32* __restrict__ vmaxSlot, u32 nSlots, SolveCtl* __restrict__ ctl) {
  // both velocity-bound slots of every scene start the solve cleared, and so does
  // the residual loop's control record (ctl: a loop follows)
  if (blockIdx.x == 0) {
    for (u32 k = threadIdx.x; k < nSlots; k += 256u) vmaxSlot[k] = 0u;
    if (ctl && threadIdx.x == 0) {
      ctl->done = 0u;
      ctl->iters = 0u;
      ctl->c0preset = 0u;
      ctl->c0sq = 0.0;
      ctl->r2 = 0.0;
    }
  }
  // each thread's first kStage list entries, written out slot by slot at the
  // end so that a store covers consecutive particles (entries found during the
  // walk land in different slots per lane)
  constexpr u32 kStage = 64;
  __shared__ u16 stage[kStage][256];
  __shared__ MatLds T;
  u32 curScene = ~0u;
  const u32 tid = threadIdx.x;
  forWork<kPersist>(W, [&](u32 scene, u32 first) {
    sceneMats(T, mats, scene, &curScene);
    const KP& P = kps[scene];
    const u32 i = first + tid;
    if (i >= P.pEnd) return;
    // kNbrWater for a solid particle of a scene with water: whether the range of
    // its candidates' tags spans a water material (a superset test: a set flag
    // only makes pbfB walk the candidates, which finds nothing to add)
    const u32 waterMask = ~T.solidMask & 0xFFu;
    const bool wantWater = waterMask != 0u && ((T.solidMask >> matOfTag(pos[i].w)) & 1u);
    float tmin = 3.0e38f, tmax = -3.0e38f;
    f3 xi = xyz(pos[i]);
    u32 n = 0, flag = 0;
    // The 9 row spans of forRowSpans, entries relative to the first candidate
    // (the lowest cell of the stencil).  The candidate walk only tests distances
    // and appends the accepted candidates to the list (candidates in batches of
    // 8, each batch's loads issued together); A is then summed over the list in
    // list order, which is the candidate order.  Summing inside the walk would
    // run the accumulation for every candidate slot in which any lane of the
    // wave accepts (nearly all of the ~216), instead of the ~30 neighbours.
    const i3 cc = cellCoord(P, xi);
    const u32 x0 = (u32)max(cc.x - 1, 0), x1 = (u32)min(cc.x + 1, P.gx - 1);
    const u32 y0 = (u32)max(cc.y - 1, 0), z0 = (u32)max(cc.z - 1, 0);
    const u32 base = cs[P.cellBase + z0 * P.gxy + y0 * (u32)P.gx + x0];
    // in registers: the list stores below may alias the scene record for all the
    // compiler knows, which would reload these from memory at every candidate
    const float supp2 = P.supp2;
    const int gy = P.gy, gz = P.gz;
    const u32 gxy = P.gxy, gx = (u32)P.gx, cellBase = P.cellBase;
    // kTrack: the scene has water, so a solid particle records the range of its
    // candidates' tags
    auto walk = [&](auto trackTag) {
      constexpr bool kTrack = decltype(trackTag)::value;
#pragma unroll 1
      for (int dz = -1; dz <= 1; dz++) {
        const int z = cc.z + dz;
        if (z < 0 || z >= gz) continue;
        const u32 plane = cellBase + (u32)z * gxy;
#pragma unroll 1
        for (int dy = -1; dy <= 1; dy++) {
          const int y = cc.y + dy;
          if (y < 0 || y >= gy) continue;
          const u32 row = plane + (u32)y * gx;
          const u32 beg = cs[row + x0];
          const u32 end = min(cs[row + x1 + 1u], beg + kRowCap);
#pragma unroll 1
          for (u32 b0 = beg; b0 < end; b0 += 8u) {
            // past the span's end the loads repeat its last candidate (no
            // predicated loads), which the test below rejects
            float4 pb[8];
#pragma unroll
            for (int j = 0; j < 8; j++) pb[j] = gat(pos, min(b0 + (u32)j, end - 1u));
#pragma unroll
            for (int j = 0; j < 8; j++) {
              const u32 c = b0 + (u32)j;
              const bool in = c < end;
              if (kTrack && wantWater && in) {
                tmin = fminf(tmin, pb[j].w);
                tmax = fmaxf(tmax, pb[j].w);
              }
              if (in && c != i && dist2(xi - xyz(pb[j])) < supp2) {
                const u32 off = c - base;
                flag |= off > kNbrOffMax ? kNbrFlag : 0u;
                const u16 e = (u16)off;
                if (n < kStage) stage[n][tid] = e;
                else if (n < K) nLst[slotAt(n, i, NT)] = e;
                n++;
              }
            }
          }
        }
      }
    };
    if (waterMask != 0u) walk(std::true_type{});
    else walk(std::false_type{});
    // A = sum_b V_b gradW(x_i - x_b) (x) (x_b - x_i) in candidate order
    f33 A = mzero<float>();
    auto addPair = [&](float4 pb, float Vb) {
      const f3 d = xi - xyz(pb);
      const f3 g = gradWfast(P, d, dist2(d));
      A = madd(A, mscale(outer(g, -d), Vb));
    };
    if (!(flag & kNbrFlag) && n <= max(K, kStage)) {
      // every neighbour is in the stage (slots below kStage) or in the list
#pragma unroll 1
      for (u32 k0 = 0; k0 < n; k0 += 8u) {
        u32 b[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const u32 k = k0 + (u32)j;
          b[j] = i;
          if (k < n) b[j] = base + (u32)(k < kStage ? stage[k][tid] : nLst[slotAt(k, i, NT)]);
        }
        float4 pb[8];
        float Vb[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          pb[j] = pos[b[j]];
          Vb[j] = vin[b[j]].w;
        }
#pragma unroll
        for (int j = 0; j < 8; j++)
          if (k0 + (u32)j < n) addPair(pb[j], Vb[j]);
      }
    } else {
      // the list cannot hold them all: the candidates again, one at a time
#pragma unroll 1
      for (int dz = -1; dz <= 1; dz++) {
        const int z = cc.z + dz;
        if (z < 0 || z >= P.gz) continue;
#pragma unroll 1
        for (int dy = -1; dy <= 1; dy++) {
          const int y = cc.y + dy;
          if (y < 0 || y >= P.gy) continue;
          const u32 row = P.cellBase + (u32)z * P.gxy + (u32)y * (u32)P.gx;
          const u32 beg = cs[row + x0];
          const u32 end = min(cs[row + x1 + 1u], beg + kRowCap);
#pragma unroll 1
          for (u32 c = beg; c < end; c++) {
            const float4 pb = pos[c];
            if (c != i && dist2(xi - xyz(pb)) < P.supp2) addPair(pb, vin[c].w);
          }
        }
      }
    }
    // the particle's own offset must fit too (the sweeps' batch tails gather i)
    flag |= (i - base) > kNbrOffMax ? kNbrFlag : 0u;
    if (wantWater && tmax >= tmin) {
      if (!(tmin >= 0.0f) || !(tmax < 8.0f)) {
        flag |= kNbrWater;  // tags outside [0, 8): no range test, walk
      } else {
        const u32 lo = (u32)tmin, hi = (u32)tmax;
        if (waterMask & ((2u << hi) - (1u << lo))) flag |= kNbrWater;
      }
    }
    // the staged slots, one group of 16 (32 bytes, in the list's interleaved
    // order: position p holds slot 16 g + 4 (p mod 4) + p / 4) at a time
    const u32 ns = min(n, min(kStage, K));  // the list may hold fewer slots than the stage
    for (u32 g = 0; 16u * g < ns; g++) {
      u32 w[8];
#pragma unroll
      for (int p = 0; p < 8; p++) {
        const u32 p0 = 2u * p, p1 = 2u * p + 1u;
        w[p] = (u32)stage[16u * g + 4u * (p0 & 3u) + (p0 >> 2)][tid] |
               ((u32)stage[16u * g + 4u * (p1 & 3u) + (p1 >> 2)][tid] << 16);
      }
      uint4* dst = (uint4*)(nLst + ((size_t)g * NT + i) * 16u);
      dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
      dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
    nHead[i] = make_uint2(base, n | flag);
    // L = pinv(A) with singular values below 0.2 dropped, in FP64.  When
    // Gershgorin's bound on A^T A puts every singular value above the cut (with
    // a margin far above FP64 rounding), pinv(A) = A^-1, formed from cofactors:
    // row k of A^-1 is (c_{k+1} x c_{k+2}) / det for the columns c of A.  Else
    // the SVD (simkernels.js:210-220).
    const d33 Ad = mcast<double>(A);
    const d3 a0 = Ad.c[0], a1 = Ad.c[1], a2 = Ad.c[2];
    const double b00 = dot(a0, a0), b11 = dot(a1, a1), b22 = dot(a2, a2);
    const double b01 = fabs(dot(a0, a1)), b02 = fabs(dot(a0, a2)), b12 = fabs(dot(a1, a2));
    const double lmin = fmin(fmin(b00 - b01 - b02, b11 - b01 - b12), b22 - b02 - b12);
    if (lmin > 0.04 * (1.0 + 1e-6)) {
      const d3 r0 = cross(a1, a2), r1 = cross(a2, a0), r2 = cross(a0, a1);
      const double id = 1.0 / dot(a0, r0);
      const d33 Ai{{d3{r0.x * id, r1.x * id, r2.x * id}, d3{r0.y * id, r1.y * id, r2.y * id},
                    d3{r0.z * id, r1.z * id, r2.z * id}}};
      stm(L, i, mcast<float>(Ai));
    } else {
      SVD3<double> s = svd3<double>(Ad);
      d3 inv = splat3(0.0);
      if (fabs(s.sig.x) > 0.2) inv.x = 1.0 / s.sig.x;
      if (fabs(s.sig.y) > 0.2) inv.y = 1.0 / s.sig.y;
      if (fabs(s.sig.z) > 0.2) inv.z = 1.0 / s.sig.z;
      stm(L, i, mcast<float>(diagMul(s.V, inv, s.U)));
    }
  });
}

// ---------------------------------------------------------------- solveA
// simkernels.js:248-309.  With part != nullptr the block also reduces the
// squared Delta-lambda numerator (the residual) and C^2 over solid particles.
template <bool kPersist>
__global__ __launch_bounds__(256) void k_solveA(const KP* __restrict__ kps, WorkTab W, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                                const float4* __restrict__ vin, M3C Fm, M3C Lm,
                                                float4* __restrict__ aux, SpP Sp, float4* __restrict__ selfDv,
                                                const u32* __restrict__ cs, NLc nl, double* __restrict__ part,
                                                const SolveCtl* __restrict__ ctl, u32 k,
                                                const float* __restrict__ hbHint, u32* __restrict__ vmaxOut) {
  if (skipIteration(ctl, k)) return;
  __shared__ MatLds T;
  u32 curScene = ~0u;
  float r2 = 0.0f, c2 = 0.0f;
  SpeedBound sb;
  forWork<kPersist>(W, [&](u32 scene, u32 first) {
    sceneMats(T, mats, scene, &curScene);
    const KP& P = kps[scene];
    const Moments mo = runMoments<kLanesA, true>(P, T, first, cs, pos, vin, nl);
    const u32 i = first + threadIdx.x;
    sb.note(scene);
    if (i >= P.pEnd) return;
    const float4 pi = pos[i];
    const f3 vi = xyz(vin[i]);
    sb.add(scene, vi);
    const u32 mi = matOfTag(pi.w);
    const bool solid = (T.solidMask >> mi) & 1u;
    const float hint = (hbHint && T.m[mi].kind == MAT_HB) ? hbHint[i] : -1.0f;
    solveAFinish(P, T, i, pi, vi, solid, mo, Fm, Lm, aux, Sp, selfDv, &r2, &c2, hint);
  });
  sb.raise(vmaxOut);
  if (part) blockReduce2((double)r2, (double)c2, part + 2 * blockIdx.x);
}

// the residual test of iteration k, run by the first block of that iteration's
// solveB (the partials of its solveA are complete once solveB starts; only
// solveA of iteration k + 1 needs the outcome): the deterministic second stage
// of the reduction, the residual history entry (out) and, with ctl, the
// stopping test and the host's mirror
struct ResidTest {
  const double* part;  // 2 doubles per block of solveA; null: no test
  u32 nb;
  double* out;
  SolveCtl* ctl;
  double tol2;
  CtlMirror* mirror;
  u32 gen;
  // the solveB of the learned first iteration: part holds the features pass's
  // partials, whose sum C^2 becomes the residual's denominator; no test
  u32 c0only;
};
__device__ __forceinline__ void residTest(const ResidTest& t, u32 k) {
  __shared__ double s_out[2];
  double a = 0.0, b = 0.0;
  for (u32 j = threadIdx.x; j < t.nb; j += 256u) {
    a += t.part[2 * j];
    b += t.part[2 * j + 1];
  }
  blockReduce2(a, b, s_out);
  __syncthreads();
  if (t.c0only) {
    if (threadIdx.x == 0) {
      t.ctl->c0sq = s_out[1];
      t.ctl->c0preset = 1u;
    }
  } else if (threadIdx.x == 0) {
    if (t.out) {
      t.out[0] = s_out[0];
      t.out[1] = s_out[1];
    }
    SolveCtl* ctl = t.ctl;
    if (ctl) {
      if (k == 0 && !ctl->c0preset) ctl->c0sq = s_out[1];
      const double c0 = ctl->c0sq;
      const u32 done = (c0 <= 0.0 || s_out[0] <= t.tol2 * c0) ? 1u : 0u;
      ctl->r2 = s_out[0];
      *reinterpret_cast<uint2*>(ctl) = make_uint2(done, k + 1u);
      if (t.mirror) {
        volatile CtlMirror* m = t.mirror;
        *reinterpret_cast<volatile uint64_t*>(m) = (uint64_t)(k + 1u) | ((uint64_t)done << 32);
        __threadfence_system();
        m->gen = t.gen;
      }
    }
  }
  __syncthreads();  // s_out is read above before a later call may write it
}

// after a loop that stopped early the velocity ping-pong has moved `iters`
// times instead of `launched`: when the two differ in parity, the final
// velocity is in `src` and is copied to `dst`, where the schedule expects it
__global__ __launch_bounds__(256) void k_parity_fix(const SolveCtl* __restrict__ ctl, u32 launched, u32 n,
                                                    const float4* __restrict__ src, float4* __restrict__ dst) {
  if (((launched - ctl->iters) & 1u) == 0u) return;
  const u32 i = xcdBlock() * 256u + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// simkernels.js:315-366
template <bool kPersist>
__global__ __launch_bounds__(256) void k_solveB(const KP* __restrict__ kps, WorkTab W, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                                const float4* __restrict__ vin, SpC Sp,
                                                const float4* __restrict__ selfDv, const u32* __restrict__ cs,
                                                NLc nl, float4* __restrict__ vout, const SolveCtl* ctl,
                                                u32 k, float4* __restrict__ xcOut, ResidTest rt,
                                                const u32* __restrict__ vmaxIn, u32* __restrict__ vmaxNext,
                                                u32 nScenes) {
  if (skipIteration(ctl, k)) return;
  if (rt.part && blockIdx.x == 0) residTest(rt, k);
  // the other slot, free since the previous solveB read it, for the next solveA
  if (blockIdx.x == 0)
    for (u32 s = threadIdx.x; s < nScenes; s += 256u) vmaxNext[s] = 0u;
  __shared__ MatLds T;
  u32 curScene = ~0u;
  forWork<kPersist>(W, [&](u32 scene, u32 first) {
    sceneMats(T, mats, scene, &curScene);
    const KP& P = kps[scene];
    const float vmax = __uint_as_float(vmaxIn[scene]);
    __shared__ float lds[SolveBAcc::kF][256];
    const SolveBAcc a = runPasses<kLanesB, SolveBAcc>(lds, first, P.pEnd, [&](u32 i, u32 q, SolveBAcc& acc) {
      const float4 pi = pos[i];
      if (!((T.solidMask >> matOfTag(pi.w)) & 1u)) return;
      const float4 vi4 = vin[i];
      if (allSolid(T)) solveBGatherQ<kLanesB, true>(P, T, i, q, pi, vi4, vmax, pos, vin, Sp, cs, nl, acc);
      else solveBGatherQ<kLanesB, false>(P, T, i, q, pi, vi4, vmax, pos, vin, Sp, cs, nl, acc);
    });
    const u32 i = first + threadIdx.x;
    if (i >= P.pEnd) return;
    const float4 pi = pos[i];
    const float4 vi4 = vin[i];
    f3 dv = splat3(0.0f);
    if ((T.solidMask >> matOfTag(pi.w)) & 1u) dv = xyz(selfDv[i]) + a.dv;
    solveBFinish(P, T, i, pi, vi4, dv, a.corr, vout, xcOut);
  });
}

// ---------------------------------------------------------------- PBF
// a candidate of pbfB: predicted position/tag, PBF multiplier
struct XL {
  float4 x;
  float l;
};
// pbfA's sums over the candidates: density, sum of gradients, sum of their squares
struct PbfAAcc {
  static constexpr int kF = 5;
  float rho;
  f3 gs;
  float g2;
  __device__ __forceinline__ static PbfAAcc zero() { return PbfAAcc{0.0f, splat3(0.0f), 0.0f}; }
  template <u32 kL> __device__ __forceinline__ void laneSum() {
    rho = xpbi::laneSum<kL>(rho);
    gs = xpbi::laneSum<kL>(gs);
    g2 = xpbi::laneSum<kL>(g2);
  }
  __device__ __forceinline__ void put(float (*lds)[256], u32 slot) const {
    lds[0][slot] = rho;
    putF3(lds, 1, slot, gs);
    lds[4][slot] = g2;
  }
  __device__ __forceinline__ static PbfAAcc get(float (*lds)[256], u32 slot) {
    return PbfAAcc{lds[0][slot], getF3(lds, 1, slot), lds[4][slot]};
  }
};
// simkernels.js:407-439; the candidate walk with four lanes per particle
template <bool kPersist>
__global__ __launch_bounds__(256) void k_pbfA(const KP* __restrict__ kps, WorkTab W, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                              const float4* __restrict__ xc, float4* __restrict__ aux,
                                              const u32* __restrict__ cs, const SolveCtl* __restrict__ ctl, u32 k) {
  if (skipIteration(ctl, k)) return;
  __shared__ MatLds T;
  __shared__ float lds[PbfAAcc::kF][256];
  __shared__ u32 qbuf[kQueue][256];
  u32 curScene = ~0u;
  forWork<kPersist>(W, [&](u32 scene, u32 first) {
    sceneMats(T, mats, scene, &curScene);
    const KP& P = kps[scene];
    const PbfAAcc a = runPasses<4u, PbfAAcc>(lds, first, P.pEnd, [&](u32 i, u32 q, PbfAAcc& acc) {
      const float4 pi = pos[i];
      const u32 mi = matOfTag(pi.w);
      if ((T.solidMask >> mi) & 1u) return;
      const float* mRho = T.mRho[mi];  // m_b / rho0 by the neighbour's material
      const f3 xci = xyz(xc[i]);
      // candidates' predicted positions (x + dt v, tag) from the iteration's solveB
      const float supp2 = P.supp2;
      const SphC K = sphConsts(P);
      forRowsQQueued(P, cs, cellCoord(P, xyz(pi)), q, qbuf, [&](u32 b) { return gat(xc, b); },
                     [&](u32 b, const float4& xb) { return b != i && dist2(xci - xyz(xb)) < supp2; },
                     [&](u32 b, const float4& xb) {
                       const f3 dxv = xci - xyz(xb);
                       const u32 kb = matOfTag(xb.w);
                       const float r = xsqrt(dist2(dxv));  // |dxv| from the support test's rounding
                       acc.rho = acc.rho + T.mass[kb] * kernelPoly6(K, r);
                       const f3 g = kernelSpikyGradR(K, dxv, r) * mRho[kb];
                       acc.gs = acc.gs + g;
                       acc.g2 = acc.g2 + dot(g, g);
                     });
    });
    const u32 i = first + threadIdx.x;
    if (i >= P.pEnd) return;
    const u32 mi = matOfTag(pos[i].w);
    if ((T.solidMask >> mi) & 1u) { aux[i].w = 0.0f; return; }
    const float rho = T.mass[mi] * P.poly6_0 + a.rho;
    const float C = wmax(rho / T.m[mi].density - 1.0f, 0.0f);
    aux[i].w = -C / (a.g2 + dot(a.gs, a.gs) + P.pbfEps);
  });
}

// simkernels.js:441-477; the candidate walk with four lanes per particle.  A
// solid particle none of whose candidates can be water (computeL's kNbrWater
// clear) gets no correction and walks nothing.
template <bool kPersist>
__global__ __launch_bounds__(256) void k_pbfB(const KP* __restrict__ kps, WorkTab W, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                              const float4* __restrict__ vin, const float4* __restrict__ xc,
                                              const float4* __restrict__ aux, const u32* __restrict__ cs,
                                              const uint2* __restrict__ head, float4* __restrict__ vout,
                                              const SolveCtl* __restrict__ ctl, u32 k) {
  if (skipIteration(ctl, k)) return;
  __shared__ MatLds T;
  __shared__ float lds[F3Acc::kF][256];
  __shared__ u32 qbuf[kQueue][256];
  u32 curScene = ~0u;
  forWork<kPersist>(W, [&](u32 scene, u32 first) {
    sceneMats(T, mats, scene, &curScene);
    const KP& P = kps[scene];
    const F3Acc a = runPasses<4u, F3Acc>(lds, first, P.pEnd, [&](u32 i, u32 q, F3Acc& acc) {
      const float4 pi = pos[i];
      const u32 mi = matOfTag(pi.w);
      const bool iAmWater = !((T.solidMask >> mi) & 1u);
      if (!iAmWater && !(head[i].y & kNbrWater)) return;
      const f3 xci = xyz(xc[i]);
      const float li = aux[i].w;
      const float supp2 = P.supp2;
      const SphC K = sphConsts(P);
      forRowsQQueued(P, cs, cellCoord(P, xyz(pi)), q, qbuf, [&](u32 b) { return XL{gat(xc, b), gatW(aux, b)}; },
                     [&](u32 b, const XL& r) {
                       const bool bWater = !((T.solidMask >> matOfTag(r.x.w)) & 1u);
                       return b != i && (iAmWater || bWater) && dist2(xci - xyz(r.x)) < supp2;
                     },
                     [&](u32 b, const XL& r) {
                       const u32 mb = matOfTag(r.x.w);
                       const bool bWater = !((T.solidMask >> mb) & 1u);
                       const f3 dxv = xci - xyz(r.x);
                       const u32 wk = bWater ? mb : mi;
                       const float rr = xsqrt(dist2(dxv));  // |dxv| from the support test's rounding
                       const float ratio = kernelPoly6(K, rr) * K.invW0;
                       const float r2 = ratio * ratio;
                       const float sCorr = -K.sCorrK * (r2 * r2);
                       acc.v = acc.v + (li + r.l + sCorr) * kernelSpikyGradR(K, dxv, rr) * T.mRho[wk][wk];
                     });
    });
    const u32 i = first + threadIdx.x;
    if (i >= P.pEnd) return;
    const float4 vi4 = vin[i];
    const float scale = (T.solidMask >> matOfTag(pos[i].w)) & 1u ? 0.5f : 1.0f;
    const f3 v = xyz(vi4) + a.v * (scale / P.dt);
    vout[i] = make_float4(v.x, v.y, v.z, vi4.w);
  });
}

// ---------------------------------------------------------------- XSPH
// simkernels.js:480-507
template <bool kPersist>
__global__ __launch_bounds__(256) void k_xsph(const KP* __restrict__ kps, WorkTab W, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                              const float4* __restrict__ vin, const u32* __restrict__ cs, NLc nl,
                                              float4* __restrict__ vout) {
  __shared__ MatLds T;
  u32 curScene = ~0u;
  forWork<kPersist>(W, [&](u32 scene, u32 first) {
    sceneMats(T, mats, scene, &curScene);
    const KP& P = kps[scene];
    __shared__ float lds[F3Acc::kF][256];
    const F3Acc a = runPasses<kLanesX, F3Acc>(lds, first, P.pEnd, [&](u32 i, u32 q, F3Acc& acc) {
      const float4 pi = pos[i];
      const f3 vi = xyz(vin[i]);
      const bool iSolid = solidTag(T.solidMask, pi.w);
      const SphC K = sphConsts(P);
      // kAll: no water in the scene, every neighbour shares the particle's phase
      auto gather = [&](auto all) {
        constexpr bool kAll = decltype(all)::value;
        forSupport<kLanesX>(P, cs, pos, nl, i, xyz(pi), q, [&](u32 b) { return PV{gat(pos, b), gat(vin, b)}; },
                    [&](u32 b, const PV& r, f3 dxv, float r2) {
                      if (kAll || solidTag(T.solidMask, r.p.w) == iSolid)
                        acc.v = acc.v + r.v.w * (xyz(r.v) - vi) * kernelWfast(K, r2);
                    });
      };
      if (allSolid(T)) gather(std::true_type{});
      else gather(std::false_type{});
    });
    const u32 i = first + threadIdx.x;
    if (i >= P.pEnd) return;
    const float4 pi = pos[i];
    const float4 vi4 = vin[i];
    float c = P.xsphC;
    if (!solidTag(T.solidMask, pi.w)) c = P.xsphC * 10.0f;
    const f3 v = xyz(vi4) + c * a.v;
    vout[i] = make_float4(v.x, v.y, v.z, vi4.w);
  });
}

// ---------------------------------------------------------------- updateF
// simkernels.js:510-525
template <bool kPersist>
__global__ __launch_bounds__(256) void k_updateF(const KP* __restrict__ kps, WorkTab W, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                                 const float4* __restrict__ vin, M3P Fm, M3C Lm,
                                                 float4* __restrict__ aux, const u32* __restrict__ cs, NLc nl,
                                                 const float* __restrict__ hbHint) {
  __shared__ MatLds T;
  u32 curScene = ~0u;
  forWork<kPersist>(W, [&](u32 scene, u32 first) {
    sceneMats(T, mats, scene, &curScene);
    const KP& P = kps[scene];
    const Moments mo = runMoments<kLanesF, false>(P, T, first, cs, pos, vin, nl);
    const u32 i = first + threadIdx.x;
    if (i >= P.pEnd) return;
    const float4 pi = pos[i];
    const u32 mi = matOfTag(pi.w);
    if (!((T.solidMask >> mi) & 1u)) return;
    const float state = aux[i].z;
    MatV m = hardenedMat(matv(T, mi), state);
    const f33 Li = ldm(Lm, i);
    const f33 gradV = mul(mo.M, transpose(Li));
    f33 Ftr = mul(madd(mident<float>(), mscale(gradV, P.dt)), ldm(Fm, i));
    SVD3<float> sv = svd3_guarded(Ftr);
    float st;
    const float hint = (hbHint && m.kind == MAT_HB) ? hbHint[i] : -1.0f;
    f3 sig = returnMap(sv.sig, m, state, P.dt, &st, hint);
    stm(Fm, i, diagMul(sv.U, safeSig(sig), sv.V));
    aux[i].z = st;
  });
}

// ---------------------------------------------------------------- integrate
// simkernels.js:528-542
template <bool kPersist>
__global__ __launch_bounds__(256) void k_integrate(const KP* __restrict__ kps, WorkTab W, float4* __restrict__ pos, float4* __restrict__ vin) {
  forWork<kPersist>(W, [&](u32 scene, u32 first) {
    const KP& P = kps[scene];
    const u32 i = first + threadIdx.x;
    if (i >= P.pEnd) return;
    float4 v4 = vin[i];
    f3 v = xyz(v4);
    float vl = len(v);
    if (vl > P.vmax) v = v * (P.vmax / vl);
    vin[i] = make_float4(v.x, v.y, v.z, v4.w);
    float4 p = pos[i];
    f3 x = xyz(p) + P.dt * v;
    x = f3{wclamp(x.x, P.dminx, P.dmaxx), wclamp(x.y, P.dminy, P.dmaxy), wclamp(x.z, P.dminz, P.dmaxz)};
    pos[i] = make_float4(x.x, x.y, x.z, p.w);
  });
}

// ---------------------------------------------------------------- learned warm start
// At lambda = 0 (before the first Jacobi iteration): the constraint of every
// solid particle as solveA evaluates it, its warm-start features, the plain
// first step d0 = (-C - gamma dt dC.v) / ((1 + gamma) denom + alpha~), and the
// unscaled Sp = GL, selfDv = sgrad that the MLP epilogue scales by lambda0.
// Feature row (16 floats): 0-2 100 log sigma of the trial F (descending),
// 3 100 tr log sigma and 4 100 |dev log sigma| after the return map,
// 5 log10(mu) - 5, 6 lambda / (lambda + 2 mu), 7 material kind / 6,
// 8 strength (DP alpha; VM, HB 100 sigmaY / mu; NACC xi / 10; snow 10 theta_c),
// 9 support neighbours / 40, 10 clamp(previous substep's lambda / d0, +-8),
// 11 alpha~ / ((1 + gamma) denom + alpha~); 12 d0 (0: no active constraint),
// 13 HB root position of the return map (NaN: none), 14 C, 15 the first step's
// numerator -C - gamma dt dC.v (its square weighs the particle in the residual).
// Also reduces sum C^2 (the residual's denominator) per block.
//
// kMlp: the learned first step in the same pass.  Each wave runs the network
// (xpbi_mlp.hpp) on its 64 particles as two MFMA passes of 32, the bf16 features
// fetched across lanes; lane l keeps the output of pass l / 32, which is its own
// particle.  Then lambda0 = clamp(y0, 0, 2) d0 (0 without an active constraint)
// goes to aux.y, the stored Sp and selfDv are lambda0 times the unscaled ones, and
// y1 is the HB bracket position of the next solveA; the feature rows are not
// written.  lambda0 is NaN-safe: fmaxf maps a NaN y0 to 0, so such a particle
// starts plain Jacobi from zero.
template <bool kPersist, bool kMlp>
__device__ __forceinline__ void featuresBody(const KP* __restrict__ kps, const WorkTab& W, const DevMat* __restrict__ mats,
                                             const float4* __restrict__ pos, const float4* __restrict__ vin, M3C Fm,
                                             M3C Lm, float4* __restrict__ aux, SpP Sp, float4* __restrict__ selfDv,
                                             const u32* __restrict__ cs, const NLc& nl,
                                             const float* __restrict__ lamPrev, float* __restrict__ feat,
                                             const MlpFwd* __restrict__ Wn, float* __restrict__ hbT,
                                             double* __restrict__ part, u32* __restrict__ vmaxOut) {
  __shared__ MatLds T;
  __shared__ std::conditional_t<kMlp, MlpFwd, char> Wl;
  if constexpr (kMlp) {
    mlp_stage_fwd(Wl, Wn);
    __syncthreads();
  }
  u32 curScene = ~0u;
  float c2 = 0.0f;
  SpeedBound sb;
  forWork<kPersist>(W, [&](u32 scene, u32 first) {
    sceneMats(T, mats, scene, &curScene);
    const KP& P = kps[scene];
    const Moments mo = runMoments<kLanesA, true>(P, T, first, cs, pos, vin, nl);
    const u32 i = first + threadIdx.x;
    sb.note(scene);
    const bool live = i < P.pEnd;
    if (!kMlp && !live) return;
    float row[16];
#pragma unroll
    for (int k = 0; k < 16; k++) row[k] = 0.0f;
    row[13] = __builtin_nanf("");
    f33 gl = mzero<float>();
    f3 sg = f3{0.0f, 0.0f, 0.0f};
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (live) {
      sb.add(scene, xyz(vin[i]));
      pi = pos[i];
      const u32 mi = matOfTag(pi.w);
      if ((T.solidMask >> mi) & 1u) {
        const float4 a = aux[i];
        const Constraint c = evalConstraint(P, T, i, pi, a, mo, Fm, Lm, -1.0f, true);
        if (c.active) {
          const MatV m = hardenedMat(matv(T, mi), a.z);
          const float g = P.kDamp;
          const float den = (1.0f + g) * c.denom + P.alphaT;
          const float d0 = (-c.C - g * P.dt * c.cdot) / den;
          const f3 lt = vlog(safeSig(c.sigTrial));
          const f3 lp = vlog(safeSig(c.sigPost));
          const float trp = lp.x + lp.y + lp.z;
          float strength = 0.0f;
          switch (m.kind) {
            case MAT_DP: strength = m.p0; break;
            case MAT_VM:
            case MAT_HB: strength = 100.0f * m.p0 / m.mu; break;
            case MAT_NACC: strength = 0.1f * m.p2; break;
            case MAT_SNOW: strength = 10.0f * m.p0; break;
            default: break;
          }
          row[0] = 100.0f * lt.x;
          row[1] = 100.0f * lt.y;
          row[2] = 100.0f * lt.z;
          row[3] = 100.0f * trp;
          row[4] = 100.0f * len(lp - splat3(trp / 3.0f));
          row[5] = log10f(m.mu) - 5.0f;
          row[6] = m.lam / (m.lam + 2.0f * m.mu);
          row[7] = (float)m.kind / 6.0f;
          row[8] = strength;
          row[9] = (float)(nl.head[i].y & ~(kNbrFlag | kNbrWater)) / 40.0f;
          row[10] = (d0 != 0.0f && lamPrev) ? wclamp(lamPrev[i] / d0, -8.0f, 8.0f) : 0.0f;
          row[11] = P.alphaT / den;
          row[12] = d0;
          row[13] = c.hbT;
          row[14] = c.C;
          row[15] = -c.C - g * P.dt * c.cdot;
          gl = c.GL;
          sg = c.sgrad;
          c2 += c.C * c.C;  // a persistent grid's thread walks several runs
        }
      }
    }
    if constexpr (!kMlp) {
      stsp(Sp, i, gl, pi);
      selfDv[i] = make_float4(sg.x, sg.y, sg.z, 0.0f);
      float4* o = (float4*)(feat + (u64)i * kFeatStride);
      o[0] = make_float4(row[0], row[1], row[2], row[3]);
      o[1] = make_float4(row[4], row[5], row[6], row[7]);
      o[2] = make_float4(row[8], row[9], row[10], row[11]);
      o[3] = make_float4(row[12], row[13], row[14], row[15]);
    } else {
      // features 0..11 as bf16 pairs; lane l of pass p takes the features 8h..8h+7
      // (h = l / 32, zero past 11) of the pass's sample l % 32, lane 32p + l % 32
      u32 xw[6];
#pragma unroll
      for (int k = 0; k < 6; k++) xw[k] = (u32)bf16_bits(row[2 * k]) | ((u32)bf16_bits(row[2 * k + 1]) << 16);
      const u32 l = threadIdx.x & 63u, hh = l >> 5;
      float y0 = 0.0f, y1 = 0.0f;
#pragma unroll 1
      for (u32 pass = 0; pass < 2u; pass++) {
        const int src = (int)(32u * pass + (l & 31u));
        u32 v[6];
#pragma unroll
        for (int k = 0; k < 6; k++) v[k] = (u32)__shfl((int)xw[k], src, 64);
        const u32 q[4] = {hh ? v[4] : v[0], hh ? v[5] : v[1], hh ? 0u : v[2], hh ? 0u : v[3]};
        bf16x8 xb;
        __builtin_memcpy(&xb, q, 16);
        u32 off = 0u;
        asm volatile("" : "+v"(off));  // keeps the weight reads in the loop: hoisted, they take ~130 registers
        MlpAct act;
        mlp_forward_b(*(const MlpFwd*)((const char*)&Wl + off), xb, act);
        if (hh == pass) {
          y0 = act.y[0];
          y1 = act.y[1];
        }
      }
      if (live) {
        const float d0 = row[12];
        const float lam = d0 == 0.0f ? 0.0f : fminf(fmaxf(y0, 0.0f), 2.0f) * d0;
        aux[i].y = lam;
        Sp.a[i] = make_float4(gl.c[0].x * lam, gl.c[0].y * lam, gl.c[0].z * lam, gl.c[1].x * lam);
        Sp.b[i] = make_float4(gl.c[1].y * lam, gl.c[1].z * lam, gl.c[2].x * lam, gl.c[2].y * lam);
        Sp.cx[i] = make_float4(pi.x, pi.y, pi.z, gl.c[2].z * lam);
        selfDv[i] = make_float4(sg.x * lam, sg.y * lam, sg.z * lam, 0.0f);
        hbT[i] = y1;
      }
    }
  });
  sb.raise(vmaxOut);
  blockReduce2(0.0, (double)c2, part + 2 * blockIdx.x);
}

template <bool kPersist>
__global__ __launch_bounds__(256) void k_features(const KP* __restrict__ kps, WorkTab W, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                                  const float4* __restrict__ vin, M3C Fm, M3C Lm,
                                                  float4* __restrict__ aux, SpP Sp, float4* __restrict__ selfDv,
                                                  const u32* __restrict__ cs, NLc nl,
                                                  const float* __restrict__ lamPrev, float* __restrict__ feat,
                                                  double* __restrict__ part, u32* __restrict__ vmaxOut) {
  featuresBody<kPersist, false>(kps, W, mats, pos, vin, Fm, Lm, aux, Sp, selfDv, cs, nl, lamPrev, feat, nullptr,
                                nullptr, part, vmaxOut);
}

// the occupancy of k_features (4 waves per SIMD, 3 for the persistent loop): the
// network's accumulators then live in VGPRs, without spills
template <bool kPersist>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kPersist ? 3 : 4))) void k_features_mlp(const KP* __restrict__ kps, WorkTab W, const DevMat* __restrict__ mats,
                                                      const float4* __restrict__ pos, const float4* __restrict__ vin,
                                                      M3C Fm, M3C Lm, float4* __restrict__ aux, SpP Sp,
                                                      float4* __restrict__ selfDv, const u32* __restrict__ cs, NLc nl,
                                                      const float* __restrict__ lamPrev, const MlpFwd* __restrict__ Wn,
                                                      float* __restrict__ hbT, double* __restrict__ part,
                                                      u32* __restrict__ vmaxOut) {
  featuresBody<kPersist, true>(kps, W, mats, pos, vin, Fm, Lm, aux, Sp, selfDv, cs, nl, lamPrev, nullptr, Wn, hbT,
                               part, vmaxOut);
}

// sum C^2 of the features pass becomes the residual's denominator
__global__ __launch_bounds__(256) void k_resid_c0(const double* __restrict__ part, u32 nb, SolveCtl* __restrict__ ctl) {
  __shared__ double s_out[2];
  double a = 0.0, b = 0.0;
  for (u32 j = threadIdx.x; j < nb; j += 256u) {
    a += part[2 * j];
    b += part[2 * j + 1];
  }
  blockReduce2(a, b, s_out);
  __syncthreads();
  if (threadIdx.x == 0) {
    ctl->c0sq = s_out[1];
    ctl->c0preset = 1u;
  }
}

// the given start (xpbi_set_lambda0): Delta-lambda = lambda0 of the particle's
// id where its constraint is active, scaling the unscaled Sp and selfDv of the
// features pass as the MLP epilogue does
__global__ __launch_bounds__(256) void k_lam0_apply(const float* __restrict__ feat, const u32* __restrict__ id,
                                                    const float* __restrict__ lam0, u32 N, float4* __restrict__ aux,
                                                    SpP Sp, float4* __restrict__ selfDv) {
  const u32 n = blockIdx.x * 256u + threadIdx.x;
  if (n >= N) return;
  const float d0 = feat[(u64)n * kFeatStride + 12];
  const float lam = d0 == 0.0f ? 0.0f : lam0[id[n]];
  aux[n].y = lam;
  const float4 A = Sp.a[n], B = Sp.b[n];
  Sp.a[n] = make_float4(A.x * lam, A.y * lam, A.z * lam, A.w * lam);
  Sp.b[n] = make_float4(B.x * lam, B.y * lam, B.z * lam, B.w * lam);
  Sp.cx[n].w = Sp.cx[n].w * lam;
  const float4 sd = selfDv[n];
  selfDv[n] = make_float4(sd.x * lam, sd.y * lam, sd.z * lam, 0.0f);
}

// training targets after the plain solve: lambda / d0 and the HB root position
__global__ __launch_bounds__(256) void k_collect_targets(const float* __restrict__ feat, const float4* __restrict__ aux,
                                                         u32 N, float* __restrict__ tgt) {
  const u32 i = blockIdx.x * 256u + threadIdx.x;
  if (i >= N) return;
  const float d0 = feat[(u64)i * kFeatStride + 12];
  tgt[2 * i] = d0 != 0.0f ? aux[i].y / d0 : __builtin_nanf("");
  tgt[2 * i + 1] = feat[(u64)i * kFeatStride + 13];
}

// ---------------------------------------------------------------- schedule
static inline u32 cdiv(u64 a, u64 b) { return (u32)((a + b - 1) / b); }
static inline M3C cm(const M3P& m) { return M3C{m.a, m.b, m.c}; }
static inline SpC spc(const SpP& m) { return SpC{m.a, m.b, m.cx}; }
// iterations the host enqueues past the last residual test it has seen
constexpr int kPollLead = 2;

// sim.js:255-266 after the reorder; ends with s->vcur on the final velocity
// the device's report for the current loop (generation gen): true once it has
// converged; returns false when it has run the residual test of iteration
// `need` (or need < 0) without converging.  Spins on the pinned mirror; past
// 0.1 ms it also lets the runtime flush the stream, past 2 s it synchronises.
static bool polledDone(Sim* s, u32 gen, int need) {
  volatile CtlMirror* m = s->mirror;
  const auto t0 = std::chrono::steady_clock::now();
  bool synced = false;
  for (u32 spin = 0;; spin++) {
    if (m->gen == gen) {
      std::atomic_thread_fence(std::memory_order_acquire);
      const uint64_t w = *reinterpret_cast<volatile const uint64_t*>(m);
      if ((u32)(w >> 32)) return true;  // done
      if ((int)(u32)w > need) return false;
    } else if (need < 0) {
      return false;
    }
    if (synced) return false;  // the stream is idle and the report is what it is
    if ((spin & 255u) == 255u) {
      const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
      if (us > 2e6) {
        hipStreamSynchronize(s->stream);
        synced = true;
      } else if (us > 100.0) {
        hipStreamQuery(s->stream);
      }
    }
  }
}

void launch_solve_stages(Sim* s, int iterations, bool water, int stop) {
  const u32 gN = s->grid;  // blocks of a particle launch (and residual partials)
  const KP* P = s->dKP;
  const WorkTab W = s->work();
  PSet& S = s->set[s->cur];
  M3P F{S.Fa, S.Fb, S.Fc};
  hipStream_t st = s->stream;
  int vi = s->vcur;
  // velocity-bound slots per scene (two sets of nScenes): the allocation-free
  // substep's host record runs one scene and keeps no scene list
  const u32 nScenes = s->sc.empty() ? 1u : (u32)s->sc.size();
  const bool loop = s->tol > 0.0f && stop != XPBI_STAGE_COMPUTE_L;
  SolveCtl* ctl = loop ? s->ctl : nullptr;
  XPBI_LAUNCH_W(k_computeL, dim3(gN), dim3(256), 0, st, P, W, (const DevMat*)s->dMats, (const float4*)S.pos,
                     (const float4*)s->vel[vi],
                     (const u32*)s->cellStart, s->L, s->nbrLst, s->nbrHead, s->nbrK, s->N, s->vmaxSlot,
                     2u * nScenes, ctl);
  const NLc nl{s->nbrHead, s->nbrLst, s->nbrK, s->N};
  // the velocity-bound slot the next solveA raises and its solveB reads
  s->vslot = 0;
  auto slotW = [&]() { return s->vmaxSlot + (size_t)s->vslot * nScenes; };
  auto slotNext = [&]() { return s->vmaxSlot + (size_t)(s->vslot ^ 1) * nScenes; };
  s->prof_mark("computeL");
  if (stop == XPBI_STAGE_COMPUTE_L) return;
  s->residIters = 0;
  const double tol2 = (double)s->tol * (double)s->tol;
  const u32 gen = loop ? ++s->solveGen : 0u;
  const float* hbHint = s->wsMode == 1 ? (const float*)s->wsHb : nullptr;
  int warm = 0;  // the learned first iteration ran
  if (s->wsMode != 0) {
    // the learned first step: the network runs inside the features pass
    const bool fused = s->wsMode == 1 && iterations > 0;
    if (fused) {
      XPBI_LAUNCH_W(k_features_mlp, dim3(gN), dim3(256), 0, st, P, W, (const DevMat*)s->dMats, (const float4*)S.pos,
                         (const float4*)s->vel[vi], cm(F), cm(s->L), S.aux, s->Sp, s->selfDv,
                         (const u32*)s->cellStart, nl, (const float*)s->lamPrev, (const MlpFwd*)s->wsPacked,
                         s->wsHb, s->residPart, slotW());
    } else {
      XPBI_LAUNCH_W(k_features, dim3(gN), dim3(256), 0, st, P, W, (const DevMat*)s->dMats, (const float4*)S.pos,
                         (const float4*)s->vel[vi], cm(F), cm(s->L), S.aux, s->Sp, s->selfDv,
                         (const u32*)s->cellStart, nl, (const float*)s->lamPrev, s->wsFeat, s->residPart, slotW());
    }
    s->prof_mark("features");
    const bool learnedFirst = (s->wsMode == 1 || s->wsMode == 3) && iterations > 0;
    // sum C^2 of the features pass: in the first block of the learned first
    // iteration's solveB, or here when there is none
    if (loop && !learnedFirst)
      hipLaunchKernelGGL(k_resid_c0, dim3(1), dim3(256), 0, st, (const double*)s->residPart, gN, ctl);
    if (learnedFirst) {
      // the first Jacobi iteration with Delta-lambda = lambda0 from the MLP (or
      // given per particle) instead of omega d0: its Sp, selfDv and lambda, then
      // solveB (and PBF)
      if (s->wsMode == 3) {
        hipLaunchKernelGGL(k_lam0_apply, dim3(cdiv(s->N, 256)), dim3(256), 0, st, (const float*)s->wsFeat,
                           (const u32*)S.id, (const float*)s->lam0ById, s->N, S.aux, s->Sp, s->selfDv);
        s->prof_mark("lam0");
      }
      XPBI_LAUNCH_W(k_solveB, dim3(gN), dim3(256), 0, st, P, W, (const DevMat*)s->dMats, (const float4*)S.pos,
                         (const float4*)s->vel[vi], spc(s->Sp), (const float4*)s->selfDv, (const u32*)s->cellStart,
                         nl, s->vel[vi ^ 1], (const SolveCtl*)nullptr, 0u, water ? s->xc : (float4*)nullptr,
                         loop ? ResidTest{(const double*)s->residPart, gN, nullptr, ctl, 0.0, nullptr, 0u, 1u}
                              : ResidTest{},
                         (const u32*)slotW(), slotNext(), nScenes);
      s->vslot ^= 1;
      s->prof_mark("solveB");
      vi ^= 1;
      if (water) {
        XPBI_LAUNCH_W(k_pbfA, dim3(gN), dim3(256), 0, st, P, W, (const DevMat*)s->dMats, (const float4*)S.pos,
                           (const float4*)s->xc, S.aux, (const u32*)s->cellStart, (const SolveCtl*)nullptr, 0u);
        s->prof_mark("pbfA");
        XPBI_LAUNCH_W(k_pbfB, dim3(gN), dim3(256), 0, st, P, W, (const DevMat*)s->dMats, (const float4*)S.pos,
                           (const float4*)s->vel[vi], (const float4*)s->xc, (const float4*)S.aux,
                           (const u32*)s->cellStart, (const uint2*)s->nbrHead, s->vel[vi ^ 1],
                           (const SolveCtl*)nullptr, 0u);
        s->prof_mark("pbfB");
        vi ^= 1;
      }
      warm = 1;
    }
  }
  int launched = 0;
  for (int k = 0; k < iterations - warm; k++) {
    if (loop && s->checkEvery > 0 && k > 0 && k % s->checkEvery == 0) {
      // stop enqueuing once the device has flagged convergence; stay at most
      // kPollLead iterations ahead of the iterations whose test the device has
      // run (the queued ones keep it busy meanwhile, about 0.25 ms each at 1M
      // particles; launches past the last iteration return at entry)
      if (polledDone(s, gen, k - kPollLead)) break;
    }
    double* out = (s->residOn && k < kMaxResidIters) ? s->residOut + 2 * k : nullptr;
    double* part = (out || loop) ? s->residPart : nullptr;
    XPBI_LAUNCH_W(k_solveA, dim3(gN), dim3(256), 0, st, P, W, (const DevMat*)s->dMats, (const float4*)S.pos,
                       (const float4*)s->vel[vi], cm(F), cm(s->L), S.aux, s->Sp, s->selfDv,
                       (const u32*)s->cellStart, nl, part, (const SolveCtl*)ctl, (u32)k, hbHint, slotW());
    s->prof_mark("solveA");
    if (out) s->residIters = k + 1;
    const ResidTest rt{(const double*)part, gN, out, ctl, tol2, loop ? s->mirror : (CtlMirror*)nullptr, gen};
    XPBI_LAUNCH_W(k_solveB, dim3(gN), dim3(256), 0, st, P, W, (const DevMat*)s->dMats, (const float4*)S.pos,
                       (const float4*)s->vel[vi], spc(s->Sp), (const float4*)s->selfDv, (const u32*)s->cellStart,
                       nl, s->vel[vi ^ 1], (const SolveCtl*)ctl, (u32)k, water ? s->xc : (float4*)nullptr, rt,
                       (const u32*)slotW(), slotNext(), nScenes);
    s->vslot ^= 1;
    s->prof_mark("solveB");
    vi ^= 1;
    if (water) {
      XPBI_LAUNCH_W(k_pbfA, dim3(gN), dim3(256), 0, st, P, W, (const DevMat*)s->dMats, (const float4*)S.pos,
                         (const float4*)s->xc, S.aux, (const u32*)s->cellStart, (const SolveCtl*)ctl, (u32)k);
      s->prof_mark("pbfA");
      XPBI_LAUNCH_W(k_pbfB, dim3(gN), dim3(256), 0, st, P, W, (const DevMat*)s->dMats, (const float4*)S.pos,
                         (const float4*)s->vel[vi], (const float4*)s->xc, (const float4*)S.aux,
                         (const u32*)s->cellStart, (const uint2*)s->nbrHead, s->vel[vi ^ 1], (const SolveCtl*)ctl,
                         (u32)k);
      s->prof_mark("pbfB");
      vi ^= 1;
    }
    launched++;
  }
  if (loop && !water && launched > 0) {
    hipLaunchKernelGGL(k_parity_fix, dim3(cdiv(s->N, 256)), dim3(256), 0, st, (const SolveCtl*)ctl, (u32)launched, s->N,
                       (const float4*)s->vel[vi ^ 1], s->vel[vi]);
    s->prof_mark("parityFix");
  }
  s->lastIters = launched;
  s->warmIters = warm;
  if (s->wsMode == 2)
    hipLaunchKernelGGL(k_collect_targets, dim3(cdiv(s->N, 256)), dim3(256), 0, st, (const float*)s->wsFeat,
                       (const float4*)S.aux, s->N, s->wsTarget);
  if (stop == XPBI_STAGE_SOLVE) { s->vcur = vi; return; }
  XPBI_LAUNCH_W(k_xsph, dim3(gN), dim3(256), 0, st, P, W, (const DevMat*)s->dMats, (const float4*)S.pos,
                     (const float4*)s->vel[vi], (const u32*)s->cellStart, nl, s->vel[vi ^ 1]);
  s->prof_mark("xsph");
  vi ^= 1;
  if (stop == XPBI_STAGE_XSPH) { s->vcur = vi; return; }
  XPBI_LAUNCH_W(k_updateF, dim3(gN), dim3(256), 0, st, P, W, (const DevMat*)s->dMats, (const float4*)S.pos,
                     (const float4*)s->vel[vi], F, cm(s->L), S.aux, (const u32*)s->cellStart, nl, hbHint);
  s->prof_mark("updateF");
  if (stop == XPBI_STAGE_UPDATE_F) { s->vcur = vi; return; }
  XPBI_LAUNCH_W(k_integrate, dim3(gN), dim3(256), 0, st, P, W, S.pos, s->vel[vi]);
  s->prof_mark("integrate");
  s->vcur = vi;
}

}  // namespace xpbi
