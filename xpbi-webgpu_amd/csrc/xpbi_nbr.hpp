// Neighbour iteration (cell rows, support-neighbour list, batched gathers) and
// the per-particle finishing code of the constraint sweeps.
#pragma once
#include <type_traits>
#include "xpbi_internal.hpp"

// Multiply-adds are fused only inside one source expression in this file, never
// across statements: a sweep's per-neighbour code rounds the same wherever it is
// inlined.
#pragma clang fp contract(on)

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

// forRowsG with four lanes per particle: lane q takes the candidates beg + q,
// beg + q + 4, ... of each row span, so that a wave-wide load reads four
// consecutive candidates of each of 16 particles; the four loads of a step are
// issued together.  use(b, record) in the lane's candidate order.
template <class Load, class Use>
__device__ __forceinline__ void forRowsQ(const KP& P, const u32* __restrict__ cs, i3 cc, u32 q, Load&& load,
                                         Use&& use) {
  forRowSpans(P, cs, cc, [&](u32 beg, u32 end) {
    for (u32 b0 = beg + q; b0 < end; b0 += 16u) {
      decltype(load(0u)) d[4];
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (b0 + 4u * (u32)j < end) d[j] = load(b0 + 4u * (u32)j);
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (b0 + 4u * (u32)j < end) use(b0 + 4u * (u32)j, d[j]);
    }
  });
}

// forRowsQ with each candidate's work split in two: test(b, rec) in the walk,
// and use(b, rec) only for the candidates the test accepts.  Each lane queues
// its accepted candidates in LDS (column threadIdx.x of qbuf, kQueue deep) and
// consumes them, reloading their records, in the lane's candidate order: when
// some lane of the wave could not queue a further step, and once the walk
// ends.  Calling use() inside the walk would run it for every candidate slot in
// which any lane of the wave accepts, which is nearly every slot.
constexpr u32 kQueue = 16;
template <class Load, class Test, class Use>
__device__ __forceinline__ void forRowsQQueued(const KP& P, const u32* __restrict__ cs, i3 cc, u32 q,
                                               u32 (*qbuf)[256], Load&& load, Test&& test, Use&& use) {
  const u32 tid = threadIdx.x;
  u32 cnt = 0;
  auto drain = [&]() {
#pragma unroll 1
    for (u32 k0 = 0; k0 < cnt; k0 += 4u) {
      u32 b[4];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = qbuf[k0 + (u32)j < cnt ? k0 + (u32)j : k0][tid];
      decltype(load(0u)) d[4];
#pragma unroll
      for (int j = 0; j < 4; j++) d[j] = load(b[j]);
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (k0 + (u32)j < cnt) use(b[j], d[j]);
    }
    cnt = 0;
  };
  forRowSpans(P, cs, cc, [&](u32 beg, u32 end) {
#pragma unroll 1
    for (u32 b0 = beg + q; b0 < end; b0 += 16u) {
      decltype(load(0u)) d[4];
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (b0 + 4u * (u32)j < end) d[j] = load(b0 + 4u * (u32)j);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const u32 b = b0 + 4u * (u32)j;
        if (b < end && test(b, d[j])) {
          qbuf[cnt][tid] = b;
          cnt++;
        }
      }
      if (__any(cnt > kQueue - 4u)) drain();  // the next step queues at most 4
    }
  });
  drain();
}

// Support-neighbour list.  Entry k of particle i is lst[k * N + i], the u16
// offset b - head[i].x of neighbour b from the particle's first candidate
// (cellStart at the stencil's lowest cell (x-1, y-1, z-1), clamped to the grid):
// the candidates of a particle span two z-layers of cells and a little more.
// head[i].y is the neighbour count; head[i].y > K (more than K neighbours, or
// kNbrFlag set for an offset past 16 bits) makes the particle's sweeps walk the
// cell rows.  Bit kNbrWater (not part of the count) is set for a solid particle
// some candidate of whose 9 row spans may be water (computeL: the candidates'
// tags span a water material); it is clear for water particles.
typedef uint16_t u16;
constexpr u32 kNbrFlag = 0x80000000u;
constexpr u32 kNbrWater = 0x40000000u;
constexpr u32 kNbrOffMax = 0xffffu;
struct NLc {
  const uint2* __restrict__ head;
  const u16* __restrict__ lst;
  u32 K, N;
};
// The list holds a particle's slots in groups of 16, interleaved for the four
// lanes of a quad (see forSupport): group g of particle i is the 32 bytes at
// lst + (g N + i) 16, and slot s = 16 g + 4 t + q (q = s mod 4) sits at position
// 4 q + t in it.  Lane q then reads its next four slots (q, q + 4, q + 8,
// q + 12 of the group) as one aligned 8-byte word, and a wave reads 16
// consecutive particles' groups as 512 contiguous bytes.  K is a multiple of 16.
__device__ __forceinline__ size_t slotAt(u32 s, u32 i, u32 N) {
  return ((size_t)(s >> 4) * N + i) * 16u + ((s & 3u) << 2) + ((s >> 2) & 3u);
}

// |d|^2 with its rounding fixed by explicit fused multiply-adds: computeL's
// acceptance test and every walk below evaluate it alike
__device__ __forceinline__ float dist2(f3 d) { return fmaf(d.z, d.z, fmaf(d.y, d.y, d.x * d.x)); }

// A walk over the particle's cell rows that yields its support neighbours one at
// a time, in the reference's candidate order: the fallback of a particle whose
// list overflowed.  next() returns the next neighbour of lane q of kL (neighbour
// number c belongs to lane c mod kL), or ~0u once the rows are exhausted.
template <u32 kL>
struct RowCursor {
  i3 cc;
  u32 x0, x1;
  int r;         // current row 0..8 (dz = r / 3 - 1, dy = r % 3 - 1); 9 = done
  u32 b, end;    // remaining candidates of the current row
  u32 cnt;       // neighbours found so far
  __device__ __forceinline__ void init(const KP& P, f3 xi) {
    cc = cellCoord(P, xi);
    x0 = (u32)max(cc.x - 1, 0);
    x1 = (u32)min(cc.x + 1, P.gx - 1);
    r = -1;
    b = end = 0;
    cnt = 0;
  }
  __device__ __forceinline__ u32 next(const KP& P, const u32* __restrict__ cs, const float4* __restrict__ pos, u32 i,
                                      f3 xi, u32 q) {
#pragma unroll 1
    for (;;) {
#pragma unroll 1
      while (b >= end) {
        if (r >= 8) return ~0u;
        r++;
        const int z = cc.z + r / 3 - 1, y = cc.y + r % 3 - 1;
        if (z < 0 || z >= P.gz || y < 0 || y >= P.gy) continue;
        const u32 row = P.cellBase + (u32)y * (u32)P.gx + (u32)z * P.gxy;
        b = cs[row + x0];
        end = min(cs[row + x1 + 1u], b + kRowCap);
      }
      const u32 c = b++;
      if (c == i) continue;
      if (dist2(xi - xyz(pos[c])) < P.supp2) {
        const bool mine = (cnt & (kL - 1u)) == q;
        cnt++;
        if (mine) return c;
      }
    }
  }
};

// Support-neighbour walk of particle i by kL = 4 or 8 lanes: lane q (0..kL-1)
// takes the neighbours numbered q, q + kL, q + 2 kL, ..., so that one wave-wide
// gather reads kL consecutive neighbours (mostly adjacent records of one cell
// row) of each of 64 / kL consecutive particles.  A lane reads the list as one
// 8-byte word per group of 16 slots: with kL = 4 the word of lane q holds its
// slots 16 g + q + 4 j (j = 0..3); with kL = 8 lane q reads the word of lane
// q mod 4 and takes its entries q / 4 and q / 4 + 2 (slots 16 g + q and
// 16 g + q + 8).  load(b) gathers a neighbour's record, whose member p is (x_b,
// tag_b); use(b, rec, dxv = x_i - x_b, |dxv|^2) consumes it, in the lane's
// order; the records of a word are loaded together.  A particle whose list
// overflowed takes its neighbours from a RowCursor instead: the same neighbours
// in the same lanes and order, through the same single copy of the loads and
// of use().
//
// forSupportLate adds a second, dependent gather: once a word's records are in,
// late(b, rec, |dxv|^2, valid) may load more of neighbour b (the contacts of
// solveB load the neighbour's velocity only for the few pairs close enough to
// touch), and use(b, rec, lateRec, dxv, |dxv|^2) consumes both.
struct NoLate {};
template <class T> __host__ __device__ const T& declRef() noexcept;  // unevaluated operands only
template <u32 kL, class Load, class Use>
__device__ __forceinline__ void forSupport(const KP& P, const u32* __restrict__ cs, const float4* __restrict__ pos,
                                           const NLc& nl, u32 i, f3 xi, u32 q, Load&& load, Use&& use);
template <u32 kL, class Load, class Late, class Use>
__device__ __forceinline__ void forSupportLate(const KP& P, const u32* __restrict__ cs,
                                               const float4* __restrict__ pos, const NLc& nl, u32 i, f3 xi, u32 q,
                                               Load&& load, Late&& late, Use&& use) {
  using Rec = decltype(load(0u));
  using LateRec = decltype(late(0u, declRef<Rec>(), 0.0f, false));
  constexpr bool kLate = !std::is_same<LateRec, NoLate>::value;
  static_assert(kL == 4u || kL == 8u, "four or eight lanes per particle");
  constexpr int kE = 16 / (int)kL;  // the lane's slots per group
  const uint2 hd = nl.head[i];
  const u32 n = hd.y & ~kNbrWater;
  const bool rows = n > nl.K;
  // list: the lane's word of group g is 8-byte word (g N + i) 4 + q mod 4; the
  // next group's word is prefetched
  const uint64_t* __restrict__ lw = (const uint64_t*)nl.lst + ((size_t)i * 4u + (q & 3u));
  const size_t gstep = (size_t)nl.N * 4u;  // 8-byte words per group
  const u32 sh = kL == 8u ? 16u * (q >> 2) : 0u;  // the lane's first entry within the word
  u32 s0 = q;                                     // the lane's first slot of the group
  uint64_t w = (!rows && s0 < n) ? lw[0] : 0ull;
  RowCursor<kL> rc;
  if (rows) rc.init(P, xi);
#pragma unroll 1
  for (u32 g = 0;; g++) {
    u32 b[kE];
    bool v[kE];
    if (!rows) {
      if (s0 >= n) break;
      const uint64_t wn = (s0 + 16u < n) ? lw[(size_t)(g + 1u) * gstep] : 0ull;
      const uint64_t ws = w >> sh;
#pragma unroll
      for (int j = 0; j < kE; j++) {
        v[j] = s0 + kL * (u32)j < n;
        b[j] = v[j] ? hd.x + (u32)((ws >> (16 * (16 / kE / 4) * j)) & 0xffffu) : i;
      }
      w = wn;
      s0 += 16u;
    } else {
#pragma unroll
      for (int j = 0; j < kE; j++) {
        const u32 c = rc.next(P, cs, pos, i, xi, q);
        v[j] = c != ~0u;
        b[j] = v[j] ? c : i;
      }
      if (!v[0]) break;
    }
    Rec d[kE];
#pragma unroll
    for (int j = 0; j < kE; j++) d[j] = load(b[j]);
    if constexpr (kLate) {
      // each record's dependent load just before its use: holding all of a
      // word's late records at once costs a wave per SIMD of occupancy
#pragma unroll
      for (int j = 0; j < kE; j++) {
        const f3 dd = xi - xyz(d[j].p);
        const float r2 = dist2(dd);
        const LateRec e = late(b[j], d[j], r2, v[j]);
        if (v[j]) use(b[j], d[j], e, dd, r2);
      }
    } else {
#pragma unroll
      for (int j = 0; j < kE; j++) {
        if (v[j]) {
          const f3 dd = xi - xyz(d[j].p);
          use(b[j], d[j], NoLate{}, dd, dist2(dd));
        }
      }
    }
    if (!v[kE - 1]) break;
  }
}
template <u32 kL, class Load, class Use>
__device__ __forceinline__ void forSupport(const KP& P, const u32* __restrict__ cs, const float4* __restrict__ pos,
                                           const NLc& nl, u32 i, f3 xi, u32 q, Load&& load, Use&& use) {
  using Rec = decltype(load(0u));
  forSupportLate<kL>(P, cs, pos, nl, i, xi, q, load, [](u32, const Rec&, float, bool) { return NoLate{}; },
                     [&](u32 b, const Rec& d, NoLate, f3 dd, float r2) { use(b, d, dd, r2); });
}

// sum over the kL lanes (4 or 8) of a particle: ((l0 + l1) + (l2 + l3)) for a
// quad, that plus the same of lanes 4..7 for eight; every lane gets the same
// value (each step adds the same pair in every lane: f32 addition commutes)
template <u32 kL>
__device__ __forceinline__ float laneSum(float x) {
  x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));  // quad_perm 1,0,3,2
  x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4E, 0xF, 0xF, false));  // quad_perm 2,3,0,1
  if (kL == 8u) x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x141, 0xF, 0xF, false));  // row_half_mirror
  return x;
}
template <u32 kL>
__device__ __forceinline__ f3 laneSum(f3 v) { return f3{laneSum<kL>(v.x), laneSum<kL>(v.y), laneSum<kL>(v.z)}; }

// The per-particle sums of a block's run of up to 256 particles, gathered by kL
// lanes per particle in kL passes of 256 / kL: in pass p the lanes kL g ..
// kL g + kL - 1 gather particle first + (256 / kL) p + g (gather(i, q, acc) adds
// lane q's share), the shares are summed, and the group's first lane parks the
// sum in LDS (lds: Acc::kF rows of 256 floats, one column per particle of the
// run).  Returns the sum of particle first + threadIdx.x, which the thread then
// finishes.  Acc provides zero(), laneSum<kL>(), put(lds, slot) and
// get(lds, slot).  Contains block barriers: every thread of the block calls it.
// The gathers run at wave priority 1 and the finish at 0: a SIMD then issues
// first for the waves whose loads keep the address unit busy, and fills the
// rest with the finishes' VALU work.
template <u32 kL, class Acc, class Gather>
__device__ __forceinline__ Acc runPasses(float (*lds)[256], u32 first, u32 pEnd, Gather&& gather) {
  constexpr u32 kPer = 256u / kL;  // particles per pass
  const u32 q = threadIdx.x & (kL - 1u), g = threadIdx.x / kL;
  __builtin_amdgcn_s_setprio(1);
#pragma unroll 1
  for (u32 pass = 0; pass < kL; pass++) {
    const u32 base = first + kPer * pass;
    if (base >= pEnd) break;
    const u32 i = base + g;
    Acc a = Acc::zero();
    if (i < pEnd) gather(i, q, a);
    a.template laneSum<kL>();
    if (q == 0u) a.put(lds, kPer * pass + g);
  }
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();
  const Acc own = Acc::get(lds, threadIdx.x);
  __syncthreads();  // the next run of a persistent block rewrites the slots
  return own;
}

// The velocity bound of solveB's contact test: each thread keeps the largest |v|
// of the particles it finished and the scenes of its runs (the same for every
// thread of the block); at the end of the kernel the block's maximum raises the
// slot of each of those scenes (an upper bound for each) by an atomic max on
// the f32 bits, which order like the values for non-negative floats.  The
// atomic is skipped when the slot already reads higher: a stale read only costs
// a redundant atomic.
struct SpeedBound {
  float s = 0.0f;
  u32 lo = ~0u, hi = 0u;
  __device__ __forceinline__ void add(u32 scene, f3 v) {
    s = fmaxf(s, len(v));
    lo = min(lo, scene);
    hi = max(hi, scene);
  }
  __device__ __forceinline__ void note(u32 scene) {
    lo = min(lo, scene);
    hi = max(hi, scene);
  }
  // every thread of the block calls it
  __device__ __forceinline__ void raise(u32* slot) {
    if (!slot) return;
    __shared__ float wmax[4];
    float m = s;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63u) == 0u) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0u && lo <= hi) {
      const u32 b = __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3])));
      for (u32 sc = lo; sc <= hi; sc++)
        if (b > *(volatile u32*)(slot + sc)) atomicMax(slot + sc, b);
    }
  }
};

__device__ __forceinline__ void putF3(float (*lds)[256], int row, u32 slot, f3 v) {
  lds[row][slot] = v.x;
  lds[row + 1][slot] = v.y;
  lds[row + 2][slot] = v.z;
}
__device__ __forceinline__ f3 getF3(float (*lds)[256], int row, u32 slot) {
  return f3{lds[row][slot], lds[row + 1][slot], lds[row + 2][slot]};
}
struct F3Acc {
  static constexpr int kF = 3;
  f3 v;
  __device__ __forceinline__ static F3Acc zero() { return F3Acc{splat3(0.0f)}; }
  template <u32 kL> __device__ __forceinline__ void laneSum() { v = xpbi::laneSum<kL>(v); }
  __device__ __forceinline__ void put(float (*lds)[256], u32 slot) const { putF3(lds, 0, slot, v); }
  __device__ __forceinline__ static F3Acc get(float (*lds)[256], u32 slot) { return F3Acc{getF3(lds, 0, slot)}; }
};

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
// the moments as a runPasses accumulator (kWithQ = false: M only)
template <bool kWithQ>
struct MomAcc {
  Moments m;
  __device__ __forceinline__ static MomAcc zero() {
    MomAcc a;
    momentsZero(a.m);
    return a;
  }
  template <u32 kL> __device__ __forceinline__ void laneSum() {
#pragma unroll
    for (int c = 0; c < 3; c++) m.M.c[c] = xpbi::laneSum<kL>(m.M.c[c]);
    if (kWithQ) {
      m.s = xpbi::laneSum<kL>(m.s);
      m.q00 = xpbi::laneSum<kL>(m.q00);
      m.q11 = xpbi::laneSum<kL>(m.q11);
      m.q22 = xpbi::laneSum<kL>(m.q22);
      m.q01 = xpbi::laneSum<kL>(m.q01);
      m.q02 = xpbi::laneSum<kL>(m.q02);
      m.q12 = xpbi::laneSum<kL>(m.q12);
    }
  }
  static constexpr int kF = kWithQ ? 18 : 9;
  __device__ __forceinline__ void put(float (*lds)[256], u32 slot) const {
#pragma unroll
    for (int c = 0; c < 3; c++) putF3(lds, 3 * c, slot, m.M.c[c]);
    if (kWithQ) {
      putF3(lds, 9, slot, m.s);
      lds[12][slot] = m.q00;
      lds[13][slot] = m.q11;
      lds[14][slot] = m.q22;
      lds[15][slot] = m.q01;
      lds[16][slot] = m.q02;
      lds[17][slot] = m.q12;
    }
  }
  __device__ __forceinline__ static MomAcc get(float (*lds)[256], u32 slot) {
    MomAcc a;
    momentsZero(a.m);
#pragma unroll
    for (int c = 0; c < 3; c++) a.m.M.c[c] = getF3(lds, 3 * c, slot);
    if (kWithQ) {
      a.m.s = getF3(lds, 9, slot);
      a.m.q00 = lds[12][slot];
      a.m.q11 = lds[13][slot];
      a.m.q22 = lds[14][slot];
      a.m.q01 = lds[15][slot];
      a.m.q02 = lds[16][slot];
      a.m.q12 = lds[17][slot];
    }
    return a;
  }
};
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
// lane q's share of particle i's moments (kAllSolid: the scene has no water,
// every neighbour contributes); the quad's shares sum to them
template <u32 kL, bool kWithQ, bool kAllSolid>
__device__ __forceinline__ void neighbourMomentsQ(const KP& P, const u32* __restrict__ cs,
                                                  const float4* __restrict__ pos, const float4* __restrict__ vin,
                                                  const NLc& nl, const float* invMass, u32 solidMask, u32 i, u32 q,
                                                  f3 xi, f3 vi, Moments& m) {
  forSupport<kL>(P, cs, pos, nl, i, xi, q, [&](u32 b) { return PV{gat(pos, b), gat(vin, b)}; },
              [&](u32 b, const PV& r, f3 dxv, float r2) {
                const float4 pb = r.p, vb = r.v;
                if (kAllSolid || solidTag(solidMask, pb.w)) {
                  momentsAdd<kWithQ>(m, gradWfast(P, dxv, r2), (xyz(vb) - vi) * vb.w, vb.w,
                                     invMass[matOfTag(pb.w)]);
                }
              });
}
// the moments of the particles of a block's run, through runPasses: each thread
// returns those of particle first + threadIdx.x (zero for water)
template <u32 kL, bool kWithQ>
__device__ __forceinline__ Moments runMoments(const KP& P, const MatLds& T, u32 first, const u32* __restrict__ cs,
                                              const float4* __restrict__ pos, const float4* __restrict__ vin,
                                              const NLc& nl) {
  __shared__ float lds[MomAcc<kWithQ>::kF][256];
  return runPasses<kL, MomAcc<kWithQ>>(lds, first, P.pEnd, [&](u32 i, u32 q, MomAcc<kWithQ>& a) {
           const float4 pi = pos[i];
           if (!((T.solidMask >> matOfTag(pi.w)) & 1u)) return;
           const f3 vi = xyz(vin[i]);
           if (allSolid(T))
             neighbourMomentsQ<kL, kWithQ, true>(P, cs, pos, vin, nl, T.invMass, T.solidMask, i, q, xyz(pi), vi,
                                                     a.m);
           else
             neighbourMomentsQ<kL, kWithQ, false>(P, cs, pos, vin, nl, T.invMass, T.solidMask, i, q, xyz(pi), vi,
                                                      a.m);
         }).m;
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
                                             const Moments& mo, M3C Fm, M3C Lm, float4* __restrict__ aux, SpP Sp,
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
      stsp(Sp, i, mscale(c.GL, dlam), pi);
      f3 sd = c.sgrad * dlam;
      selfDv[i] = make_float4(sd.x, sd.y, sd.z, 0.0f);
      wrote = true;
      *r2 += num * num;
      *c2 += c.C * c.C;
    }
  }
  if (!wrote) {
    stsp(Sp, i, mzero<float>(), pi);
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

// a neighbour's records of solveB's constraint gather: p = (x_b, the ninth float
// of Sp_b), a and b = Sp_b's first eight floats, and with water the tag of b
// a neighbour's records of solveB: all solid, p = (x_b, the ninth float of Sp_b)
// and a, b = Sp_b's first eight floats; with water also the neighbour's
// position/tag p4 and velocity up front
struct SpGather {
  float4 p, a, b;
};
struct SpGatherW {
  float4 p, v, a, b;
  float c;
};
// an all-solid contact candidate's velocity and tag, loaded only when the pair
// may touch
struct SpLate {
  float4 v;
  float tag;
  bool near;
};

// simkernels.js:338-355 for one neighbour: the constraint gather
// dv += V_i S_b gradW(x_b - x_i) and, when `near`, the 0.75 h contact on
// predicted positions (xb, vb: the neighbour's position and velocity, tagB its
// material tag).  All solid, the contact test |x_i + dt v_i - x_b - dt v_b| <
// 0.75 h cannot pass unless |x_i - x_b| < 0.75 h + dt (|v_i| + |v_b|): past the
// bound rc2 on |dxv|^2 (with the largest |v_b| of the velocities read) the
// pair's velocity is not loaded and the test is not evaluated.  The bound
// carries a margin above the rounding of the f32 test, so the skipped pairs are
// exactly those the test rejects.
__device__ __forceinline__ void solveBPair(const SphC& P, float Vi, const float* wc_i, f3 xci, f3 dxv, float r2,
                                           const f33& Spb, bool near, f3 xb, float4 vb, float tagB, f3* dv,
                                           f3* corr) {
  const f3 dxp = -dxv;  // x_b - x_i
  *dv = *dv + Vi * mul(Spb, gradWfast(P, dxp, r2));
  if (near) {
    const f3 xcb = xb + P.dt * xyz(vb);
    const f3 dxc = xci - xcb;
    const float d2 = dist2(dxc);
    const float minD = P.minD;
    if (d2 < minD * minD && d2 > 1e-18f) {
      float idist = xrsq(d2);
      float dist = d2 * idist;
      float w = wc_i[matOfTag(tagB)];
      *corr = *corr + dxc * (idist * ((minD - dist) * w));
    }
  }
}

// |dxv|^2 past which a pair of particle i cannot touch in solveB's contact test
// (see solveBPair): vmax bounds |v_b| over the velocities read; the margin covers
// the f32 rounding of x + dt v and of the test's difference (a few ulp of the
// coordinates) and of the bound itself
__device__ __forceinline__ float contactBound2(const KP& P, f3 xi, f3 vi, float vmax) {
  const float reach = P.dt * (len(vi) + vmax);
  const float xmax = fmaxf(fmaxf(fabsf(xi.x), fabsf(xi.y)), fabsf(xi.z)) + P.support + reach;
  const float rc = (P.minD + reach) * (1.0f + 1e-5f) + 4e-6f * xmax;
  return rc * rc;
}

// lane q's share of the constraint gather dv += V_i Sp_b gradW(x_b - x_i) and
// of the contact correction (selfDv is added by the owner of the particle)
struct SolveBAcc {
  static constexpr int kF = 6;
  f3 dv, corr;
  __device__ __forceinline__ static SolveBAcc zero() { return SolveBAcc{splat3(0.0f), splat3(0.0f)}; }
  template <u32 kL> __device__ __forceinline__ void laneSum() {
    dv = xpbi::laneSum<kL>(dv);
    corr = xpbi::laneSum<kL>(corr);
  }
  __device__ __forceinline__ void put(float (*lds)[256], u32 slot) const {
    putF3(lds, 0, slot, dv);
    putF3(lds, 3, slot, corr);
  }
  __device__ __forceinline__ static SolveBAcc get(float (*lds)[256], u32 slot) {
    return SolveBAcc{getF3(lds, 0, slot), getF3(lds, 3, slot)};
  }
};
// All solid: three 16-byte records per neighbour, (x_b, Sp_b's ninth float) and
// Sp_b's first eight floats; a contact candidate then loads its velocity and
// tag.  With water (whose pairs the solid test skips, and whose clumps put many
// pairs within contact reach): position/tag, velocity and Sp up front.
template <u32 kL, bool kAllSolid>
__device__ __forceinline__ void solveBGatherQ(const KP& P, const MatLds& T, u32 i, u32 q, float4 pi, float4 vi4,
                                              float vmax, const float4* __restrict__ pos,
                                              const float4* __restrict__ vin, SpC Sp, const u32* __restrict__ cs,
                                              const NLc& nl, SolveBAcc& a) {
  const float Vi = vi4.w;
  const f3 xi = xyz(pi);
  const f3 xci = xi + P.dt * xyz(vi4);
  const float* wc_i = T.wc[matOfTag(pi.w)];
  const SphC K = sphConsts(P);
  if constexpr (kAllSolid) {
    const float rc2 = contactBound2(P, xi, xyz(vi4), vmax);
    forSupportLate<kL>(P, cs, pos, nl, i, xi, q,
        [&](u32 b) { return SpGather{gat(Sp.cx, b), gat(Sp.a, b), gat(Sp.b, b)}; },
        [&](u32 b, const SpGather& g, float r2, bool valid) {
          SpLate e{make_float4(0.0f, 0.0f, 0.0f, 0.0f), 0.0f, false};
          if (valid && !(r2 >= rc2)) {
            e.v = gat(vin, b);
            e.tag = gatW(pos, b);
            e.near = true;
          }
          return e;
        },
        [&](u32 b, const SpGather& g, const SpLate& e, f3 dxv, float r2) {
          const f33 Spb{{f3{g.a.x, g.a.y, g.a.z}, f3{g.a.w, g.b.x, g.b.y}, f3{g.b.z, g.b.w, g.p.w}}};
          solveBPair(K, Vi, wc_i, xci, dxv, r2, Spb, e.near, xyz(g.p), e.v, e.tag, &a.dv, &a.corr);
        });
  } else {
    const float* __restrict__ c9 = (const float*)Sp.cx + 3;
    forSupport<kL>(P, cs, pos, nl, i, xi, q,
        [&](u32 b) { return SpGatherW{pos[b], vin[b], Sp.a[b], Sp.b[b], c9[4 * (size_t)b]}; },
        [&](u32 b, const SpGatherW& g, f3 dxv, float r2) {
          if (!solidTag(T.solidMask, g.p.w)) return;
          const f33 Spb{{f3{g.a.x, g.a.y, g.a.z}, f3{g.a.w, g.b.x, g.b.y}, f3{g.b.z, g.b.w, g.c}}};
          solveBPair(K, Vi, wc_i, xci, dxv, r2, Spb, true, xyz(g.p), g.v, g.p.w, &a.dv, &a.corr);
        });
  }
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
