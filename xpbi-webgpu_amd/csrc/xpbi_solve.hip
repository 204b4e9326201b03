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
      u32 row = (u32)y * (u32)P.gx + (u32)z * P.gxy;
      u32 beg = cs[row + x0];
      u32 end = min(cs[row + x1 + 1u], beg + kRowCap);
#pragma unroll 1
      for (u32 b = beg; b < end; b++) fn(b);
    }
  }
}

struct NLc {
  const u32* __restrict__ idx;  // idx[k * N + i]
  const u32* __restrict__ cnt;
  u32 K, N;
};

// fn(b, pos[b], dxv = x_i - x_b, |dxv|^2) for each support neighbour of i.  A
// particle whose list overflowed (cnt > K) walks the cell rows instead.
template <class Fn>
__device__ __forceinline__ void forSupport(const KP& P, const u32* __restrict__ cs, const float4* __restrict__ pos,
                                           const NLc& nl, u32 i, f3 xi, Fn&& fn) {
  const u32 n = nl.cnt[i];
  if (n <= nl.K) {
    const u32* __restrict__ p = nl.idx + i;
#pragma unroll 2
    for (u32 k = 0; k < n; k++) {
      const u32 b = p[(size_t)k * nl.N];
      const float4 pb = pos[b];
      const f3 d = xi - xyz(pb);
      fn(b, pb, d, dot(d, d));
    }
  } else {
    forRows(P, cs, cellCoord(P, xi), [&](u32 b) {
      if (b != i) {
        const float4 pb = pos[b];
        const f3 d = xi - xyz(pb);
        const float r2 = dot(d, d);
        if (r2 < P.supp2) fn(b, pb, d, r2);
      }
    });
  }
}

#ifndef XPBI_GATHER_BATCH
#define XPBI_GATHER_BATCH 8
#endif
constexpr int kBatch = XPBI_GATHER_BATCH;

// As forSupport, with the neighbour gathers issued in batches of kBatch
// independent loads (and the next batch's indices prefetched) so that several
// cache misses are in flight per lane: load(b) gathers the per-neighbour data,
// use(b, pos[b], dxv, |dxv|^2, data) consumes it, in list order.
template <class Load, class Use>
__device__ __forceinline__ void forSupportG(const KP& P, const u32* __restrict__ cs, const float4* __restrict__ pos,
                                            const NLc& nl, u32 i, f3 xi, Load&& load, Use&& use) {
  const u32 n = nl.cnt[i];
  if (n <= nl.K) {
    const u32* __restrict__ p = nl.idx + i;
    u32 bn[kBatch];
#pragma unroll
    for (int j = 0; j < kBatch; j++) bn[j] = ((u32)j < n) ? p[(size_t)j * nl.N] : i;
    for (u32 k = 0; k < n; k += kBatch) {
      u32 b[kBatch];
      float4 pb[kBatch];
      decltype(load(0u)) d[kBatch];
#pragma unroll
      for (int j = 0; j < kBatch; j++) {
        b[j] = bn[j];
        pb[j] = pos[b[j]];
        d[j] = load(b[j]);
      }
#pragma unroll
      for (int j = 0; j < kBatch; j++) {
        u32 kk = k + kBatch + (u32)j;
        bn[j] = (kk < n) ? p[(size_t)kk * nl.N] : i;
      }
#pragma unroll
      for (int j = 0; j < kBatch; j++) {
        if (k + (u32)j < n) {
          const f3 dd = xi - xyz(pb[j]);
          use(b[j], pb[j], dd, dot(dd, dd), d[j]);
        }
      }
    }
  } else {
    forRows(P, cs, cellCoord(P, xi), [&](u32 b) {
      if (b != i) {
        const float4 pb = pos[b];
        const f3 d = xi - xyz(pb);
        const float r2 = dot(d, d);
        if (r2 < P.supp2) use(b, pb, d, r2, load(b));
      }
    });
  }
}

// material tables staged in LDS at kernel entry
struct MatLds {
  DevMat m[8];
  float mass[8];
  float invMass[8];
  u32 solidMask;  // bit k set when material k is not water
};
__device__ __forceinline__ void loadMatLds(MatLds& t, const DevMat* __restrict__ mats) {
  const u32 tid = threadIdx.x;
  if (tid < 128) ((float*)t.m)[tid] = ((const float*)mats)[tid];
  if (tid < 8) {
    t.mass[tid] = mats[tid].mass;
    t.invMass[tid] = 1.0f / mats[tid].mass;
  }
  if (tid == 0) {
    u32 mask = 0;
    for (int k = 0; k < 8; k++)
      if (mats[k].kind != MAT_WATER) mask |= 1u << k;
    t.solidMask = mask;
  }
  __syncthreads();
}
__device__ __forceinline__ MatV matv(const MatLds& t, u32 k) {
  const DevMat& r = t.m[k];
  return MatV{r.mu, r.lam, r.density, r.kind, r.p0, r.p1, r.p2, r.p3, r.boundFriction, r.mass};
}
__device__ __forceinline__ bool solidTag(u32 mask, float w) { return (mask >> matOfTag(w)) & 1u; }

// ---------------------------------------------------------------- computeL
// simkernels.js:189-221.  A is accumulated in f32; the SVD pseudo-inverse with
// absolute tolerance 0.2 runs in FP64.  The same candidate walk records the
// support-neighbour list of the substep.
__global__ __launch_bounds__(256) void k_computeL(KP P, const float4* __restrict__ pos, const float4* __restrict__ vin,
                                                  const u32* __restrict__ cs, M3P L, u32* __restrict__ nIdx,
                                                  u32* __restrict__ nCnt, u32 K) {
  u32 i = xcdBlock() * 256u + threadIdx.x;
  if (i >= P.N) return;
  f3 xi = xyz(pos[i]);
  f33 A = mzero<float>();
  u32 n = 0;
  forRows(P, cs, cellCoord(P, xi), [&](u32 b) {
    if (b != i) {
      f3 dxv = xi - xyz(pos[b]);
      float r2 = dot(dxv, dxv);
      if (r2 < P.supp2) {
        f3 g = gradWfast(P, dxv, r2);
        f3 r = -dxv;
        A = madd(A, mscale(outer(g, r), vin[b].w));
        if (n < K) nIdx[(size_t)n * P.N + i] = b;
        n++;
      }
    }
  });
  nCnt[i] = n;
  SVD3<double> s = svd3<double>(mcast<double>(A));
  d3 inv = splat3(0.0);
  if (fabs(s.sig.x) > 0.2) inv.x = 1.0 / s.sig.x;
  if (fabs(s.sig.y) > 0.2) inv.y = 1.0 / s.sig.y;
  if (fabs(s.sig.z) > 0.2) inv.z = 1.0 / s.sig.z;
  stm(L, i, mcast<float>(diagMul(s.V, inv, s.U)));
}

// ---------------------------------------------------------------- velocity gradient
// simkernels.js:224-245.  With g_b = gradW(x_i - x_b) and w_b = V_b (v_b - v_i),
//   grad v = sum_b w_b (x) (L_i g_b) = (sum_b w_b (x) g_b) L_i^T = M L_i^T,
// so the sweep accumulates the first moment M and L_i is applied once.
struct Moments {
  f33 M;    // sum_b V_b (v_b - v_i) (x) g_b
  f3 s;     // sum_b V_b g_b
  float q00, q11, q22, q01, q02, q12;  // sum_b (V_b^2 / m_b) g_b g_b^T
};
template <bool kWithQ>
__device__ __forceinline__ Moments neighbourMoments(const KP& P, const u32* __restrict__ cs,
                                                    const float4* __restrict__ pos, const float4* __restrict__ vin,
                                                    const NLc& nl, const float* invMass, u32 solidMask, u32 i,
                                                    f3 xi, f3 vi) {
  Moments m;
  m.M = mzero<float>();
  m.s = splat3(0.0f);
  m.q00 = m.q11 = m.q22 = m.q01 = m.q02 = m.q12 = 0.0f;
  forSupportG(P, cs, pos, nl, i, xi, [&](u32 b) { return vin[b]; },
              [&](u32 b, float4 pb, f3 dxv, float r2, float4 vb) {
                if (solidTag(solidMask, pb.w)) {
                  const f3 g = gradWfast(P, dxv, r2);
                  const f3 w = (xyz(vb) - vi) * vb.w;
                  m.M = madd(m.M, outer(w, g));
                  if (kWithQ) {
                    m.s = m.s + g * vb.w;
                    const float c = vb.w * vb.w * invMass[matOfTag(pb.w)];
                    const f3 cg = g * c;
                    m.q00 += cg.x * g.x;
                    m.q11 += cg.y * g.y;
                    m.q22 += cg.z * g.z;
                    m.q01 += cg.x * g.y;
                    m.q02 += cg.x * g.z;
                    m.q12 += cg.y * g.z;
                  }
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

// ---------------------------------------------------------------- solveA
// simkernels.js:248-309.  With part != nullptr the block also reduces the
// squared Delta-lambda numerator (the residual) and C^2 over solid particles.
__global__ __launch_bounds__(256) void k_solveA(KP P, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                                const float4* __restrict__ vin, M3C Fm, M3C Lm,
                                                float4* __restrict__ aux, M3P Sp, float4* __restrict__ selfDv,
                                                const u32* __restrict__ cs, NLc nl, double* __restrict__ part) {
  __shared__ MatLds T;
  loadMatLds(T, mats);
  const u32 i = xcdBlock() * 256u + threadIdx.x;
  float r2 = 0.0f, c2 = 0.0f;
  if (i < P.N) {
    const float4 pi = pos[i];
    const u32 mi = matOfTag(pi.w);
    bool wrote = false;
    if ((T.solidMask >> mi) & 1u) {
      const float4 a = aux[i];
      MatV m = hardenedMat(matv(T, mi), a.z);
      const f3 xi = xyz(pi);
      const f33 Fn = ldm(Fm, i);
      const f33 Li = ldm(Lm, i);
      const float4 vi4 = vin[i];
      const f3 vi = xyz(vi4);
      const Moments mo = neighbourMoments<true>(P, cs, pos, vin, nl, T.invMass, T.solidMask, i, xi, vi);
      const f33 gradV = mul(mo.M, transpose(Li));
      f33 Ftr = mul(madd(mident<float>(), mscale(gradV, P.dt)), Fn);
      SVD3<float> sv = svd3_guarded(Ftr);
      float st;
      f3 sig = returnMap(sv.sig, m, a.z, P.dt, &st);
      float psi = stvkPsi(sig, m.mu, m.lam);
      float C = sqrtf(wmax(2.0f * psi, 0.0f));
      if (C >= 1e-7f) {
        f3 dpsi = stvkDPsi(sig, m.mu, m.lam);
        f33 Pk = diagMul(sv.U, dpsi / C, sv.V);
        f33 GL = mul(mul(Pk, transpose(Fn)), Li);
        // the reference's second sweep (simkernels.js:284-299) from the moments:
        //   g'_b = V_b GL g_b, sgrad = -sum g'_b = -GL s,
        //   sum |g'_b|^2 / m_b = (GL^T GL) : Q,  cdot = sum g'_b.(v_b - v_i) = GL : M
        const f3 sgrad = -mul(GL, mo.s);
        const float h00 = dot(GL.c[0], GL.c[0]), h11 = dot(GL.c[1], GL.c[1]), h22 = dot(GL.c[2], GL.c[2]);
        const float h01 = dot(GL.c[0], GL.c[1]), h02 = dot(GL.c[0], GL.c[2]), h12 = dot(GL.c[1], GL.c[2]);
        float denom = h00 * mo.q00 + h11 * mo.q11 + h22 * mo.q22 + 2.0f * (h01 * mo.q01 + h02 * mo.q02 + h12 * mo.q12);
        denom = denom + dot(sgrad, sgrad) / m.mass;
        const float cdot = frob(GL, mo.M);
        const float g = P.kDamp;
        const float num = -C - P.alphaT * a.y - g * P.dt * cdot;
        const float dlam = P.omega * num / ((1.0f + g) * denom + P.alphaT);
        aux[i].y = a.y + dlam;
        stm(Sp, i, mscale(GL, dlam));
        f3 sd = sgrad * dlam;
        selfDv[i] = make_float4(sd.x, sd.y, sd.z, 0.0f);
        wrote = true;
        r2 = num * num;
        c2 = C * C;
      }
    }
    if (!wrote) {
      stm(Sp, i, mzero<float>());
      selfDv[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  }
  if (part) blockReduce2((double)r2, (double)c2, part + 2 * blockIdx.x);
}

// deterministic second stage of the residual reduction
__global__ __launch_bounds__(256) void k_resid_final(const double* __restrict__ part, u32 nb, double* __restrict__ out) {
  double a = 0.0, b = 0.0;
  for (u32 k = threadIdx.x; k < nb; k += 256u) {
    a += part[2 * k];
    b += part[2 * k + 1];
  }
  blockReduce2(a, b, out);
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
  float4 v, a, b;
  float c;
};

// simkernels.js:315-366
__global__ __launch_bounds__(256) void k_solveB(KP P, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                                const float4* __restrict__ vin, M3C Sp,
                                                const float4* __restrict__ selfDv, const u32* __restrict__ cs,
                                                NLc nl, float4* __restrict__ vout) {
  __shared__ MatLds T;
  loadMatLds(T, mats);
  const u32 i = xcdBlock() * 256u + threadIdx.x;
  if (i >= P.N) return;
  const float4 pi = pos[i];
  const u32 mi = matOfTag(pi.w);
  const f3 xi = xyz(pi);
  const float4 vi4 = vin[i];
  const float mass_i = T.mass[mi];
  f3 dv = splat3(0.0f), corr = splat3(0.0f);
  if ((T.solidMask >> mi) & 1u) {
    float4 sd = selfDv[i];
    dv = xyz(sd);
    const float Vi = vi4.w;
    const f3 xci = xi + P.dt * xyz(vi4);
    const float minD = P.minD;
    forSupportG(P, cs, pos, nl, i, xi, [&](u32 b) { return SpGather{vin[b], Sp.a[b], Sp.b[b], Sp.c[b]}; },
                [&](u32 b, float4 pb4, f3 dxv, float r2, const SpGather& g) {
      if (solidTag(T.solidMask, pb4.w)) {
        f3 pb = xyz(pb4);
        f3 dxp = -dxv;  // x_b - x_i
        const f33 Spb{{f3{g.a.x, g.a.y, g.a.z}, f3{g.a.w, g.b.x, g.b.y}, f3{g.b.z, g.b.w, g.c}}};
        dv = dv + Vi * mul(Spb, gradWfast(P, dxp, r2));
        f3 xcb = pb + P.dt * xyz(g.v);
        f3 dxc = xci - xcb;
        float d2 = dot(dxc, dxc);
        if (d2 < minD * minD && d2 > 1e-18f) {
          float idist = xrsq(d2);
          float dist = d2 * idist;
          float mb = T.mass[matOfTag(pb4.w)];
          float w = mb / (mass_i + mb);
          corr = corr + dxc * (idist * ((minD - dist) * w));
        }
      }
    });
  }
  f3 dvv = dv / (mass_i * P.dt);
  float dvl = len(dvv);
  if (dvl > P.dvCap) dvv = dvv * (P.dvCap / dvl);
  f3 v = xyz(vi4) + dvv + corr * (P.contactOmega / P.dt);
  v = applyBoundary(P, xi + P.dt * v, v, xi, T.m[mi].boundFriction);
  vout[i] = make_float4(v.x, v.y, v.z, vi4.w);
}

// ---------------------------------------------------------------- PBF
// simkernels.js:407-439
__global__ __launch_bounds__(256) void k_pbfA(KP P, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                              const float4* __restrict__ vin, float4* __restrict__ aux,
                                              const u32* __restrict__ cs) {
  __shared__ MatLds T;
  loadMatLds(T, mats);
  const u32 i = xcdBlock() * 256u + threadIdx.x;
  if (i >= P.N) return;
  const float4 pi = pos[i];
  const u32 mi = matOfTag(pi.w);
  if ((T.solidMask >> mi) & 1u) { aux[i].w = 0.0f; return; }
  const float rho0 = T.m[mi].density;
  const f3 xn = xyz(pi);
  const f3 xci = xn + P.dt * xyz(vin[i]);
  float rho = T.mass[mi] * P.poly6_0;
  f3 gradSum = splat3(0.0f);
  float grad2 = 0.0f;
  forRows(P, cs, cellCoord(P, xn), [&](u32 b) {
    if (b != i) {
      float4 pb = pos[b];
      f3 xcb = xyz(pb) + P.dt * xyz(vin[b]);
      f3 dxv = xci - xcb;
      if (dot(dxv, dxv) < P.supp2) {
        float mb = T.mass[matOfTag(pb.w)];
        rho = rho + mb * kernelPoly6(P, len(dxv));
        f3 g = kernelSpikyGrad(P, dxv) * (mb / rho0);
        gradSum = gradSum + g;
        grad2 = grad2 + dot(g, g);
      }
    }
  });
  float C = wmax(rho / rho0 - 1.0f, 0.0f);
  aux[i].w = -C / (grad2 + dot(gradSum, gradSum) + P.pbfEps);
}

// simkernels.js:441-477
__global__ __launch_bounds__(256) void k_pbfB(KP P, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                              const float4* __restrict__ vin, const float4* __restrict__ aux,
                                              const u32* __restrict__ cs, float4* __restrict__ vout) {
  __shared__ MatLds T;
  loadMatLds(T, mats);
  const u32 i = xcdBlock() * 256u + threadIdx.x;
  if (i >= P.N) return;
  const float4 pi = pos[i];
  const u32 mi = matOfTag(pi.w);
  const f3 xn = xyz(pi);
  const float4 vi4 = vin[i];
  const f3 xci = xn + P.dt * xyz(vi4);
  f3 dp = splat3(0.0f);
  const float li = aux[i].w;
  const bool iAmWater = !((T.solidMask >> mi) & 1u);
  forRows(P, cs, cellCoord(P, xn), [&](u32 b) {
    if (b != i) {
      float4 pb = pos[b];
      u32 mb = matOfTag(pb.w);
      bool bWater = !((T.solidMask >> mb) & 1u);
      if (iAmWater || bWater) {
        f3 xcb = xyz(pb) + P.dt * xyz(vin[b]);
        f3 dxv = xci - xcb;
        if (dot(dxv, dxv) < P.supp2) {
          u32 wk = bWater ? mb : mi;
          float lb = aux[b].w;
          float w = kernelPoly6(P, len(dxv));
          float ratio = w / wmax(P.w0, 1e-12f);
          float r2 = ratio * ratio;
          float sCorr = -P.sCorrK * (r2 * r2);
          dp = dp + (li + lb + sCorr) * kernelSpikyGrad(P, dxv) * (T.m[wk].mass / T.m[wk].density);
        }
      }
    }
  });
  float scale = iAmWater ? 1.0f : 0.5f;
  f3 v = xyz(vi4) + dp * (scale / P.dt);
  vout[i] = make_float4(v.x, v.y, v.z, vi4.w);
}

// ---------------------------------------------------------------- XSPH
// simkernels.js:480-507
__global__ __launch_bounds__(256) void k_xsph(KP P, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                              const float4* __restrict__ vin, const u32* __restrict__ cs, NLc nl,
                                              float4* __restrict__ vout) {
  __shared__ MatLds T;
  loadMatLds(T, mats);
  const u32 i = xcdBlock() * 256u + threadIdx.x;
  if (i >= P.N) return;
  const float4 pi = pos[i];
  const f3 xi = xyz(pi);
  const float4 vi4 = vin[i];
  const f3 vi = xyz(vi4);
  const bool iSolid = solidTag(T.solidMask, pi.w);
  f3 acc = splat3(0.0f);
  forSupportG(P, cs, pos, nl, i, xi, [&](u32 b) { return vin[b]; },
              [&](u32 b, float4 pb, f3 dxv, float r2, float4 vb) {
                if (solidTag(T.solidMask, pb.w) == iSolid) acc = acc + vb.w * (xyz(vb) - vi) * kernelWfast(P, r2);
              });
  float c = P.xsphC;
  if (!iSolid) c = P.xsphC * 10.0f;
  f3 v = vi + c * acc;
  vout[i] = make_float4(v.x, v.y, v.z, vi4.w);
}

// ---------------------------------------------------------------- updateF
// simkernels.js:510-525
__global__ __launch_bounds__(256) void k_updateF(KP P, const DevMat* __restrict__ mats, const float4* __restrict__ pos,
                                                 const float4* __restrict__ vin, M3P Fm, M3C Lm,
                                                 float4* __restrict__ aux, const u32* __restrict__ cs, NLc nl) {
  __shared__ MatLds T;
  loadMatLds(T, mats);
  const u32 i = xcdBlock() * 256u + threadIdx.x;
  if (i >= P.N) return;
  const float4 pi = pos[i];
  const u32 mi = matOfTag(pi.w);
  if (!((T.solidMask >> mi) & 1u)) return;
  const float state = aux[i].z;
  MatV m = hardenedMat(matv(T, mi), state);
  const f3 xi = xyz(pi);
  const f33 Li = ldm(Lm, i);
  const Moments mo = neighbourMoments<false>(P, cs, pos, vin, nl, T.invMass, T.solidMask, i, xi, xyz(vin[i]));
  const f33 gradV = mul(mo.M, transpose(Li));
  f33 Ftr = mul(madd(mident<float>(), mscale(gradV, P.dt)), ldm(Fm, i));
  SVD3<float> sv = svd3_guarded(Ftr);
  float st;
  f3 sig = returnMap(sv.sig, m, state, P.dt, &st);
  stm(Fm, i, diagMul(sv.U, safeSig(sig), sv.V));
  aux[i].z = st;
}

// ---------------------------------------------------------------- integrate
// simkernels.js:528-542
__global__ __launch_bounds__(256) void k_integrate(KP P, float4* __restrict__ pos, float4* __restrict__ vin) {
  const u32 i = xcdBlock() * 256u + threadIdx.x;
  if (i >= P.N) return;
  float4 v4 = vin[i];
  f3 v = xyz(v4);
  float vl = len(v);
  if (vl > P.vmax) v = v * (P.vmax / vl);
  vin[i] = make_float4(v.x, v.y, v.z, v4.w);
  float4 p = pos[i];
  f3 x = xyz(p) + P.dt * v;
  x = f3{wclamp(x.x, P.dminx, P.dmaxx), wclamp(x.y, P.dminy, P.dmaxy), wclamp(x.z, P.dminz, P.dmaxz)};
  pos[i] = make_float4(x.x, x.y, x.z, p.w);
}

// ---------------------------------------------------------------- schedule
static inline u32 cdiv(u64 a, u64 b) { return (u32)((a + b - 1) / b); }
static inline M3C cm(const M3P& m) { return M3C{m.a, m.b, m.c}; }

// sim.js:255-266 after the reorder; ends with s->vcur on the final velocity
void launch_solve_stages(Sim* s, int iterations, bool water, int stop) {
  const u32 gN = cdiv(s->N, 256);
  const KP& P = s->kp;
  PSet& S = s->set[s->cur];
  M3P F{S.Fa, S.Fb, S.Fc};
  hipStream_t st = s->stream;
  int vi = s->vcur;
  hipLaunchKernelGGL(k_computeL, dim3(gN), dim3(256), 0, st, P, (const float4*)S.pos, (const float4*)s->vel[vi],
                     (const u32*)s->cellStart, s->L, s->nbrIdx, s->nbrCnt, s->nbrK);
  const NLc nl{s->nbrIdx, s->nbrCnt, s->nbrK, s->N};
  s->prof_mark("computeL");
  if (stop == XPBI_STAGE_COMPUTE_L) return;
  s->residIters = 0;
  for (int k = 0; k < iterations; k++) {
    double* part = (s->residOn && k < kMaxResidIters) ? s->residPart : nullptr;
    hipLaunchKernelGGL(k_solveA, dim3(gN), dim3(256), 0, st, P, (const DevMat*)s->dMats, (const float4*)S.pos,
                       (const float4*)s->vel[vi], cm(F), cm(s->L), S.aux, s->Sp, s->selfDv,
                       (const u32*)s->cellStart, nl, part);
    s->prof_mark("solveA");
    if (part) {
      hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(256), 0, st, (const double*)part, gN, s->residOut + 2 * k);
      s->residIters = k + 1;
    }
    hipLaunchKernelGGL(k_solveB, dim3(gN), dim3(256), 0, st, P, (const DevMat*)s->dMats, (const float4*)S.pos,
                       (const float4*)s->vel[vi], cm(s->Sp), (const float4*)s->selfDv, (const u32*)s->cellStart,
                       nl, s->vel[vi ^ 1]);
    s->prof_mark("solveB");
    vi ^= 1;
    if (water) {
      hipLaunchKernelGGL(k_pbfA, dim3(gN), dim3(256), 0, st, P, (const DevMat*)s->dMats, (const float4*)S.pos,
                         (const float4*)s->vel[vi], S.aux, (const u32*)s->cellStart);
      s->prof_mark("pbfA");
      hipLaunchKernelGGL(k_pbfB, dim3(gN), dim3(256), 0, st, P, (const DevMat*)s->dMats, (const float4*)S.pos,
                         (const float4*)s->vel[vi], (const float4*)S.aux, (const u32*)s->cellStart, s->vel[vi ^ 1]);
      s->prof_mark("pbfB");
      vi ^= 1;
    }
  }
  if (stop == XPBI_STAGE_SOLVE) { s->vcur = vi; return; }
  hipLaunchKernelGGL(k_xsph, dim3(gN), dim3(256), 0, st, P, (const DevMat*)s->dMats, (const float4*)S.pos,
                     (const float4*)s->vel[vi], (const u32*)s->cellStart, nl, s->vel[vi ^ 1]);
  s->prof_mark("xsph");
  vi ^= 1;
  if (stop == XPBI_STAGE_XSPH) { s->vcur = vi; return; }
  hipLaunchKernelGGL(k_updateF, dim3(gN), dim3(256), 0, st, P, (const DevMat*)s->dMats, (const float4*)S.pos,
                     (const float4*)s->vel[vi], F, cm(s->L), S.aux, (const u32*)s->cellStart, nl);
  s->prof_mark("updateF");
  if (stop == XPBI_STAGE_UPDATE_F) { s->vcur = vi; return; }
  hipLaunchKernelGGL(k_integrate, dim3(gN), dim3(256), 0, st, P, S.pos, s->vel[vi]);
  s->prof_mark("integrate");
  s->vcur = vi;
}

}  // namespace xpbi
