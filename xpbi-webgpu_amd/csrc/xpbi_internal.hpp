// Host-side state of one simulation and the launchers shared by the .hip files.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "xpbi_dev.hpp"
#include "xpbi.h"

namespace xpbi {

typedef uint64_t u64;

// particle arrays that are gathered into cell order every substep (two copies:
// the reorder reads one set and writes the other)
struct PSet {
  float4* pos;
  float4* Fa;
  float4* Fb;
  float* Fc;
  float4* aux;
  u32* id;
};
struct PSetC {
  const float4* pos;
  const float4* Fa;
  const float4* Fb;
  const float* Fc;
  const float4* aux;
  const u32* id;
};

// device-side state of the residual-driven Jacobi loop.  The iteration-k
// residual is r_k^2 = sum num^2 / sum C0^2 over solid particles, where num is
// the numerator of Delta-lambda / omega at the start of iteration k and C0 the
// constraints of iteration 0.  Once r_k <= tol, iteration k is the last one
// applied: later launches of the loop exit at once.
struct SolveCtl {
  // done and iters form one 8-byte word, written and read whole: the test of
  // iteration k runs in the first block of that iteration's solveB while the
  // kernel's other blocks read the pair
  u32 done;
  u32 iters;  // Jacobi iterations applied
  u32 c0preset;  // c0sq was set by the warm-start features pass
  u32 pad;
  double c0sq;
  double r2;  // last residual numerator^2
};

// what the device reports to the host after each residual test, written by the
// device straight into pinned host memory, so that the host polls it while the
// stream keeps running (no copy, no synchronisation)
struct CtlMirror {
  // iters and done form one 8-byte word, written and read whole, so that the
  // host never pairs a new iteration count with the flag of an earlier test
  u32 iters;  // iterations whose residual test has run
  u32 done;   // converged
  u32 gen;    // solve the fields above belong to (written last)
  u32 pad;
};

// iteration k of the loop is skipped once an earlier iteration was the last:
// the stopping iteration itself (iters == k + 1) still runs its solveB / PBF
__device__ __forceinline__ bool skipIteration(const SolveCtl* ctl, u32 k) {
  if (!ctl) return false;
  const uint2 di = *reinterpret_cast<const uint2*>(ctl);  // (done, iters)
  return di.x && di.y <= k;
}

constexpr int kMaxProfileMarks = 4096;
constexpr int kMaxResidIters = 4096;

// one scene of a simulation (a batch holds several; a single scene is a batch of one)
struct Scene {
  u32 N = 0, C = 0;
  u32 pBase = 0, cellBase = 0;  // its particles / cells in the shared arrays and cell table
  int dims[3] = {0, 0, 0};
  float origin[3] = {0, 0, 0};
  float cell = 0;
  double h = 0;
  bool haveParams = false, haveMats = false;
  xpbi_params params{};
  xpbi_material mats[8]{};
  KP kp{};
};

struct Sim {
  int dev = 0;
  hipStream_t stream = nullptr;
  u32 N = 0, C = 0;  // all scenes
  std::vector<Scene> sc;
  KP* dKP = nullptr;       // per scene
  DevMat* dMats = nullptr;  // 8 per scene
  uint2* dTab = nullptr;   // work table: (scene, first particle) per 256-particle run
  u32 nb = 0;              // work table entries
  u32 grid = 0;            // blocks of a particle launch (nb, or a persistent grid)
  bool persist = false;    // grid < nb: the kernels' persistent instantiation

  // captured two-substep graph and the configuration it was captured with
  hipGraphExec_t graphExec = nullptr;
  hipGraph_t graph = nullptr;
  long long graphKey[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  WorkTab work() const { return WorkTab{dTab, nb}; }

  PSet set[2]{};
  int cur = 0;
  float4* vel[2] = {nullptr, nullptr};
  int vcur = 0;  // index of the current velocity ("velA")
  M3P L{};
  SpP Sp{};
  float4* selfDv = nullptr;
  // per scene, max |v| over the velocities an iteration's solveB reads (as f32
  // bits), two slots: solveA (or the warm-start features pass) raises slot
  // vslot, solveB reads it and clears the other for the next iteration; computeL
  // clears both
  u32* vmaxSlot = nullptr;
  int vslot = 0;
  u32* cellStart = nullptr;  // C + 1
  u32* count = nullptr;      // C
  u32* keys[2] = {nullptr, nullptr};
  u32* vals[2] = {nullptr, nullptr};
  u32* perm = nullptr;        // points into vals[]
  u32* sortedKeys = nullptr;  // points into keys[]
  u32* scanPart = nullptr;
  // support-neighbour list of the current substep, built by computeL:
  // support-neighbour list: nbrLst[k * N + i] for k < nbrHead[i].y <= nbrK, each
  // a u16 offset from the base nbrHead[i].x (nbrHead[i].y > nbrK: the sweeps use
  // the cell rows)
  uint16_t* nbrLst = nullptr;
  uint2* nbrHead = nullptr;
  // 512 B per particle.  Clumping PBF water in the 1M dam breach has 110 neighbours
  // at q99 and up to ~200; a particle past nbrK walks all its cell rows (800+
  // candidates in the densest cells), and its whole wave waits for it
  u32 nbrK = 256;
  static constexpr u32 kMaxNbrSlots = 1024;
  float4* xc = nullptr;  // water scenes: predicted positions x + dt v (w = tag) after each solveB
  bool sorted = false;  // a substep has run (cellStart/perm valid)

  // residual monitor
  bool residOn = false;
  double* residPart = nullptr;  // 2 * blocks
  double* residOut = nullptr;   // 2 * kMaxResidIters
  int residIters = 0;

  // residual-driven loop (tol > 0): `iterations` becomes a cap
  float tol = 0.0f;
  int checkEvery = 4;       // host polls the done flag every checkEvery iterations (0: never)
  SolveCtl* ctl = nullptr;  // device
  SolveCtl* ctlHost = nullptr;  // pinned copy of ctl
  CtlMirror* mirror = nullptr;  // pinned, written by the residual test in solveB
  u32 solveGen = 0;             // generation of the current residual loop
  int lastIters = 0;        // fixed-iteration count of the last substep (tol == 0)
  int warmIters = 0;        // 1 when the last substep's first iteration was the learned one

  // learned warm start (mode 1) and training-pair collection (mode 2)
  int wsMode = 0;
  void* wsPacked = nullptr;  // packed MLP weights (device)
  float* wsFeat = nullptr;   // N x 16 feature rows of the current substep
  float* wsHb = nullptr;     // N: HB bracket hint (mode 1)
  float* wsTarget = nullptr; // N x 2 training targets (mode 2)
  float* lam0ById = nullptr; // N: given lambda0 per persistent id (mode 3)
  float* lamPrev = nullptr;  // N: previous substep's final lambda, in the current order

  // profiling
  bool profiling = false;
  std::vector<hipEvent_t> events;
  std::vector<std::string> labels;
  int nmarks = 0;
  void prof_mark(const char* label) {
    if (!profiling || nmarks + 1 >= (int)events.size()) return;
    hipEventRecord(events[nmarks + 1], stream);
    labels[nmarks] = label;
    nmarks++;
  }

  std::vector<void*> allocs;
  std::string err;
};

// a work-table kernel in the instantiation that matches s->grid (s in scope)
#define XPBI_LAUNCH_W(K, ...)                                   \
  do {                                                          \
    if (s->persist) hipLaunchKernelGGL(K<true>, __VA_ARGS__);   \
    else hipLaunchKernelGGL(K<false>, __VA_ARGS__);             \
  } while (0)

// launchers
void launch_sort(Sim* s);
void launch_reorder(Sim* s);
void launch_exclusive_scan(Sim* s, const u32* in, u32 n, u32* out, bool write_total, u32 total);
size_t scan_partials_needed(u64 n);
void launch_solve_stages(Sim* s, int iterations, bool water, int stop);
void launch_copy(Sim* s, const float4* src, float4* dst, u64 n);
void mlp_pack_host(const float* params, void* packed);
size_t mlp_packed_bytes();

}  // namespace xpbi

struct xpbi_sim {
  xpbi::Sim s;
};
