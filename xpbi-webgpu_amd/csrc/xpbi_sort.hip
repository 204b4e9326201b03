// Grid hash, stable counting sort and cell-order reorder (gfx950).
//
// Replaces the reference's clearBuffer + countParticles + scanBlocks +
// scanApply + scatter + stagePVA/applyPVA/stageF/applyF
// (web/js/sim.js:223-254, web/js/shaders/simkernels.js:78-186).
//
//  * k_hash_count: cell id per particle (common.js:97-103) and the per-cell
//    histogram (atomicAdd on integer counts: the counts are order independent).
//  * exclusive scan of the counts -> cellStart[0..C] with cellStart[C] = N; the
//    scan is a wave64 shuffle scan inside reduce-then-scan blocks, replacing the
//    reference's O(NB^2) serial block-sum loop (simkernels.js:117-121).
//  * the permutation is the stable counting sort: ordered by cell, and inside a
//    cell by ascending original index -- deterministic, unlike the reference's
//    atomicSub scatter (simkernels.js:138).  The count's atomics hand each
//    particle a slot in its cell (in whatever order they land); k_cell_scatter
//    puts the particle there, and k_cell_rank moves it to its stable place,
//    cellStart + the number of the cell's particles with a smaller index.
//  * k_reorder gathers every particle array into cell order in one pass and
//    fuses the per-particle work of scatter: V^n = V0 clamp(det F, 0.1, 10),
//    v += dt g, lambda = 0 (simkernels.js:142-146).
#include "xpbi_internal.hpp"

namespace xpbi {

constexpr int SCAN_T = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_T * SCAN_ITEMS;  // 4096 elements per block

// p[0..n) = v, four items per thread in one 16-B store (p is 16-B aligned)
__global__ __launch_bounds__(256) void k_fill_u32(u32* __restrict__ p, u32 n, u32 v) {
  const u32 q = blockIdx.x * 256u + threadIdx.x;
  if (4u * q + 4u <= n) {
    ((uint4*)p)[q] = make_uint4(v, v, v, v);
  } else {
    for (u32 i = 4u * q; i < n; i++) p[i] = v;
  }
}

// countParticles (simkernels.js:79-84): the cell id of each particle and the
// per-cell counts, and each particle's slot among its cell's particles in the
// order the atomics landed
template <bool kPersist>
__global__ __launch_bounds__(256) void k_hash_count(const KP* __restrict__ kps, WorkTab W, const float4* __restrict__ pos, u32* __restrict__ keys,
                                                    u32* __restrict__ count, u32* __restrict__ slot) {
  forWork<kPersist>(W, [&](u32 scene, u32 first) {
    const KP& P = kps[scene];
    const u32 i = first + threadIdx.x;
    const bool valid = i < P.pEnd;
    u32 c = ~0u;
    if (valid) {
      float4 p = pos[i];
      c = cellIndex(P, cellCoord(P, f3{p.x, p.y, p.z}));
      keys[i] = c;
    }
    // the array is in the previous substep's cell order, so consecutive
    // particles mostly share a cell: one atomic per run of equal cells within
    // the wave (a run ends where the next lane's cell differs), adding the run's
    // length.  The counts are the same integers in any order.
    const u32 lane = threadIdx.x & 63u;
    const u32 prev = __shfl_up(c, 1u, 64);
    const bool head = lane == 0u || prev != c;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = (2ull << lane) - 1ull;  // lanes 0..lane (lane 63: all)
    u32 old = 0u;
    if (valid && head) {
      const uint64_t after = heads & ~upto;  // heads of the lanes above
      const u32 next = after ? (u32)__ffsll((unsigned long long)after) - 1u : 64u;
      old = atomicAdd(&count[c], next - lane);
    }
    // the run's slots follow the count its head's atomic found
    const u32 hl = 63u - (u32)__clzll((long long)(heads & upto));
    const u32 base = __shfl(old, hl, 64);
    if (valid) slot[i] = base + (lane - hl);
  });
}

// each particle into its slot of its cell (cellStart[c] + slot): cell order,
// the order inside a cell that of the atomics
__global__ __launch_bounds__(256) void k_cell_scatter(const u32* __restrict__ keys, const u32* __restrict__ slot,
                                                      const u32* __restrict__ cellStart, u32 n,
                                                      u32* __restrict__ ckeys, u32* __restrict__ cidx) {
  const u32 i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const u32 c = keys[i];
  const u32 d = cellStart[c] + slot[i];
  ckeys[d] = c;
  cidx[d] = i;
}

// the stable place of the particle in slot s: its cell's start plus the number
// of the cell's particles with a smaller original index.  The lanes of a wave
// mostly share a cell and walk the same entries; the work is the sum over cells
// of their squared counts.
__global__ __launch_bounds__(256) void k_cell_rank(const u32* __restrict__ ckeys, const u32* __restrict__ cidx,
                                                   const u32* __restrict__ cellStart, u32 n,
                                                   u32* __restrict__ skeys, u32* __restrict__ perm) {
  const u32 s = blockIdx.x * 256u + threadIdx.x;
  if (s >= n) return;
  const u32 c = ckeys[s], i = cidx[s];
  const u32 b = cellStart[c], e = cellStart[c + 1];
  u32 r = 0;
#pragma unroll 4
  for (u32 k = b; k < e; k++) r += cidx[k] < i ? 1u : 0u;
  skeys[b + r] = c;
  perm[b + r] = i;
}

// ---------------------------------------------------------------- scan
__device__ __forceinline__ u32 wave_incl_scan(u32 v, u32 lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    u32 t = __shfl_up(v, o, 64);
    if (lane >= (u32)o) v += t;
  }
  return v;
}
// block-wide exclusive scan of one value per thread (256 threads); returns the
// exclusive prefix and writes the block total to *total
__device__ __forceinline__ u32 block_excl_scan(u32 v, u32* total) {
  __shared__ u32 s_w[4];
  u32 lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  u32 inc = wave_incl_scan(v, lane);
  if (lane == 63u) s_w[w] = inc;
  __syncthreads();
  u32 off = 0, tot = 0;
#pragma unroll
  for (u32 k = 0; k < 4; k++) {
    u32 s = s_w[k];
    if (k < w) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// the tile's sum: loads of 16 B each, consecutive lanes on consecutive
// addresses (the sum does not need each thread's items to be contiguous)
__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const u32* __restrict__ in, u32 n, u32* __restrict__ part) {
  const u32 tile = blockIdx.x * (u32)SCAN_TILE;
  u32 s = 0;
  if (tile + (u32)SCAN_TILE <= n) {
    const uint4* p = (const uint4*)(in + tile);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; k++) {
      const uint4 q = p[k * SCAN_T + threadIdx.x];
      s += q.x + q.y + q.z + q.w;
    }
  } else {
    for (u32 idx = tile + threadIdx.x; idx < n; idx += SCAN_T) s += in[idx];
  }
  u32 tot;
  (void)block_excl_scan(s, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}
// single block: exclusive scan of the nb block totals, in place
__global__ __launch_bounds__(SCAN_T) void k_scan_partials(u32* __restrict__ part, u32 nb) {
  u32 carry = 0;
  for (u32 base = 0; base < nb; base += SCAN_T) {
    u32 idx = base + threadIdx.x;
    u32 v = idx < nb ? part[idx] : 0u;
    u32 tot;
    u32 ex = block_excl_scan(v, &tot);
    if (idx < nb) part[idx] = carry + ex;
    carry += tot;
  }
}
// out[i] = exclusive prefix of in (in may alias out); out[n] = total if wtot.
// Each thread scans its 16 consecutive items, moved as four 16-B loads and
// stores (in and out are 16-B aligned).  The block's offset is part[blockIdx.x]
// after k_scan_partials, or with sumPart the sum of the tile totals before it
// (the raw part of k_scan_reduce; integer sums, the same in any order).
__global__ __launch_bounds__(SCAN_T) void k_scan_down(const u32* in, u32 n, const u32* __restrict__ part, u32* out,
                                                     int wtot, u32 total_if_known, int sumPart) {
  const u32 base = blockIdx.x * (u32)SCAN_TILE + threadIdx.x * (u32)SCAN_ITEMS;
  const bool full = base + (u32)SCAN_ITEMS <= n;
  u32 v[SCAN_ITEMS];
  if (full) {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; k++) {
      const uint4 q = ((const uint4*)(in + base))[k];
      v[4 * k] = q.x;
      v[4 * k + 1] = q.y;
      v[4 * k + 2] = q.z;
      v[4 * k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) v[k] = base + k < n ? in[base + k] : 0u;
  }
  u32 s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) s += v[k];
  u32 off;
  if (sumPart) {
    u32 acc = 0;
    for (u32 j = threadIdx.x; j < blockIdx.x; j += SCAN_T) acc += part[j];
    (void)block_excl_scan(acc, &off);
  } else {
    off = part[blockIdx.x];
  }
  u32 tot;
  u32 run = off + block_excl_scan(s, &tot);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    const u32 x = v[k];
    v[k] = run;
    run += x;
  }
  if (full) {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; k++)
      ((uint4*)(out + base))[k] = make_uint4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++)
      if (base + k < n) out[base + k] = v[k];
  }
  if (wtot && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_T - 1) out[n] = total_if_known;
}

// ---------------------------------------------------------------- reorder
template <bool kPersist>
__global__ __launch_bounds__(256) void k_reorder(const KP* __restrict__ kps, WorkTab W, const u32* __restrict__ perm, PSetC src,
                                                 const float4* __restrict__ vsrc, PSet dst, float4* __restrict__ vdst,
                                                 float* __restrict__ lamPrev) {
  forWork<kPersist>(W, [&](u32 scene, u32 first) {
    const KP& P = kps[scene];
    const u32 j = first + threadIdx.x;
    if (j >= P.pEnd) return;
    u32 i = perm[j];
    float4 p = src.pos[i];
    float4 v = vsrc[i];
    float4 a = src.aux[i];
    f33 F = ldm(src.Fa, src.Fb, src.Fc, i);
    u32 id = src.id[i];
    float Vn = P.V0 * wclamp(det3(F), 0.1f, 10.0f);
    v.x = v.x + P.dt * P.grx;
    v.y = v.y + P.dt * P.gry;
    v.z = v.z + P.dt * P.grz;
    v.w = Vn;
    a.x = Vn;
    if (lamPrev) lamPrev[j] = a.y;  // the previous substep's final lambda, for the warm start
    a.y = 0.0f;
    dst.pos[j] = p;
    vdst[j] = v;
    dst.aux[j] = a;
    stm(M3P{dst.Fa, dst.Fb, dst.Fc}, j, F);
    dst.id[j] = id;
  });
}

// plain float4 copy (bandwidth ceiling / counter calibration)
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, u64 n) {
  u64 i = (u64)blockIdx.x * 256u + threadIdx.x;
  u64 stride = (u64)gridDim.x * 256u;
  for (; i < n; i += stride) dst[i] = src[i];
}

// ---------------------------------------------------------------- host side
static inline u32 cdiv(u64 a, u64 b) { return (u32)((a + b - 1) / b); }

// in and out 16-B aligned (device allocations and the scratch carve are 256-B aligned)
void launch_exclusive_scan(Sim* s, const u32* in, u32 n, u32* out, bool write_total, u32 total) {
  u32 nb = cdiv(n, SCAN_TILE);
  if (nb == 0) nb = 1;
  // up to kSumPart tiles each down-sweep block sums the totals before it (at
  // most kSumPart reads per block); beyond, a single block scans them first
  constexpr u32 kSumPart = 2048;
  const int sumPart = nb <= kSumPart ? 1 : 0;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_T), 0, s->stream, in, n, s->scanPart);
  if (!sumPart) hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(SCAN_T), 0, s->stream, s->scanPart, nb);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_T), 0, s->stream, in, n, (const u32*)s->scanPart, out,
                     write_total ? 1 : 0, total, sumPart);
}

size_t scan_partials_needed(u64 n) { return (size_t)cdiv(n, SCAN_TILE) + 1; }

void launch_sort(Sim* s) {
  const u32 N = s->N, C = s->C;
  hipLaunchKernelGGL(k_fill_u32, dim3(cdiv(cdiv(C, 4), 256)), dim3(256), 0, s->stream, s->count, C, 0u);
  s->prof_mark("clear");
  XPBI_LAUNCH_W(k_hash_count, dim3(s->grid), dim3(256), 0, s->stream, (const KP*)s->dKP, s->work(),
                     (const float4*)s->set[s->cur].pos, s->keys[0], s->count, s->vals[0]);
  s->prof_mark("countParticles");
  launch_exclusive_scan(s, s->count, C, s->cellStart, true, N);
  s->prof_mark("scanCells");
  if (N > 0) {
    hipLaunchKernelGGL(k_cell_scatter, dim3(cdiv(N, 256)), dim3(256), 0, s->stream, (const u32*)s->keys[0],
                       (const u32*)s->vals[0], (const u32*)s->cellStart, N, s->keys[1], s->vals[1]);
    hipLaunchKernelGGL(k_cell_rank, dim3(cdiv(N, 256)), dim3(256), 0, s->stream, (const u32*)s->keys[1],
                       (const u32*)s->vals[1], (const u32*)s->cellStart, N, s->keys[0], s->vals[0]);
  }
  s->perm = s->vals[0];
  s->sortedKeys = s->keys[0];
  s->prof_mark("cellSort");
}

void launch_reorder(Sim* s) {
  int src = s->cur, dst = s->cur ^ 1;
  PSet& S = s->set[src];
  PSetC sc{S.pos, S.Fa, S.Fb, S.Fc, S.aux, S.id};
  XPBI_LAUNCH_W(k_reorder, dim3(s->grid), dim3(256), 0, s->stream, (const KP*)s->dKP, s->work(),
                     (const u32*)s->perm, sc,
                     (const float4*)s->vel[s->vcur], s->set[dst], s->vel[s->vcur ^ 1],
                     s->wsMode != 0 ? s->lamPrev : (float*)nullptr);
  s->cur = dst;
  s->vcur ^= 1;
  s->prof_mark("reorder");
}

void launch_copy(Sim* s, const float4* src, float4* dst, u64 n) {
  u32 blocks = (u32)((n + 255) / 256);
  if (blocks > 8192u) blocks = 8192u;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, s->stream, src, dst, n);
}

}  // namespace xpbi
